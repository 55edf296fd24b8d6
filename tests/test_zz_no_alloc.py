"""bem_op_potential_noalloc (device memory from a static array instead of
hipMalloc) against bem_op_potential, through the C-ABI: two calls in a row on
the same operator, the second with fewer points and an element mask, for a
complex (Helmholtz single layer, SurfaceL2 0) and a real operator with a
three-component evaluator (Laplace hypersingular, H1 1); then a call whose
sources exceed the array (icosphere level 5 at intorder 12: 49 points x 20,480
triangles x 80 bytes = 80 MB > 64 MiB) is rejected with an error and leaves
the output untouched.  Both entry points run the same kernels with the same
source splits, so the values agree to rounding (bar 1e-13 relative)."""
import ctypes

import numpy as np
import pytest
import torch

import ngbem_amd as nb
from ngbem_amd import _capi

pytestmark = pytest.mark.gpu


def _call(fn, op, coef, P, N, mask, out):
    L = _capi.lib()
    stream = torch.cuda.current_stream().cuda_stream
    return getattr(L, fn)(op._h, ctypes.c_void_p(coef.data_ptr()), ctypes.c_void_p(P.data_ptr()),
                          None if N is None else ctypes.c_void_p(N.data_ptr()), P.shape[0],
                          None if mask is None else mask.ctypes.data_as(ctypes.c_void_p),
                          ctypes.c_void_p(out.data_ptr()), _capi.C128 if op.is_complex else _capi.F64,
                          ctypes.c_void_p(stream))


def _points(n, seed, gpu):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n, 3))
    r = rng.uniform(1.3, 3.0, n)[:, None]
    P = torch.from_numpy(d / np.linalg.norm(d, axis=1, keepdims=True) * r).to(gpu).contiguous()
    N = torch.from_numpy(rng.standard_normal((n, 3))).to(gpu).contiguous()
    return P, N


@pytest.mark.parametrize("kind", ["HSL-P0", "LHS-P1"])
def test_potential_noalloc_matches_potential(gpu, kind):
    mesh = nb.icosphere(3)
    kw = dict(intorder=4, leafsize=32, eps=1e-8)
    if kind == "HSL-P0":
        sp = nb.SurfaceL2(mesh)
        op = nb.HelmholtzSingleLayerPotentialOperator(sp, sp, kappa=4.0, **kw)
    else:
        sp = nb.H1(mesh)
        op = nb.HypersingularOperator(sp, **kw)
    dt = torch.complex128 if op.is_complex else torch.float64
    rng = np.random.default_rng(7)
    c = rng.standard_normal(sp.ndof) + (1j * rng.standard_normal(sp.ndof) if op.is_complex else 0)
    coef = torch.from_numpy(c).to(gpu).to(dt).contiguous()
    dim = _capi.lib().bem_op_potential_dim(op._h)
    print(f"case {kind}: ndof {sp.ndof}, dim {dim}", flush=True)
    half = np.zeros(mesh.nt, dtype=np.uint8)
    half[: mesh.nt // 2] = 1
    for npts, mask, seed in ((3000, None, 1), (400, half, 2)):
        P, N = _points(npts, seed, gpu)
        ref = torch.zeros(npts, dim, dtype=dt, device=gpu)
        got = torch.full((npts, dim), 7.0, dtype=dt, device=gpu)
        _capi.check(_call("bem_op_potential", op, coef, P, N, mask, ref), "bem_op_potential")
        _capi.check(_call("bem_op_potential_noalloc", op, coef, P, N, mask, got), "bem_op_potential_noalloc")
        torch.cuda.synchronize()
        scale = ref.abs().max().item()
        assert scale > 0
        err = (got - ref).abs().max().item()
        print(f"  {npts} points, mask {mask is not None}: max |noalloc - potential| / max |potential| {err / scale:.1e}",
              flush=True)
        assert err <= 1e-13 * scale


def test_potential_noalloc_rejects_oversize(gpu):
    mesh = nb.icosphere(5)
    sp = nb.SurfaceL2(mesh)
    op = nb.HelmholtzSingleLayerPotentialOperator(sp, sp, kappa=2.0, intorder=12, leafsize=64, eps=1e-6)
    coef = torch.ones(sp.ndof, dtype=torch.complex128, device=gpu)
    P, N = _points(100, 3, gpu)
    out = torch.full((100, 1), 7.0, dtype=torch.complex128, device=gpu)
    rc = _call("bem_op_potential_noalloc", op, coef, P, N, None, out)
    msg = _capi.lib().bem_last_error().decode()
    print(f"rc {rc}: {msg}", flush=True)
    assert rc != 0 and "scratch" in msg and "bytes" in msg
    torch.cuda.synchronize()
    assert torch.all(out == 7.0)
    # the array is not left in a bad state: a call that fits still works
    small = nb.SurfaceL2(nb.icosphere(2))
    op2 = nb.HelmholtzSingleLayerPotentialOperator(small, small, kappa=2.0, intorder=4, leafsize=32, eps=1e-6)
    c2 = torch.ones(small.ndof, dtype=torch.complex128, device=gpu)
    ref = torch.zeros(100, 1, dtype=torch.complex128, device=gpu)
    _capi.check(_call("bem_op_potential", op2, c2, P, N, None, ref), "bem_op_potential")
    _capi.check(_call("bem_op_potential_noalloc", op2, c2, P, N, None, out), "bem_op_potential_noalloc")
    torch.cuda.synchronize()
    assert (out - ref).abs().max().item() <= 1e-13 * ref.abs().max().item()
