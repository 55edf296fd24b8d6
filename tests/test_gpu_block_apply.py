"""Block product of many right-hand sides (MFMA path) against the vector path
and the oracle, through the C-ABI.

A batch X of k row vectors goes through bem_op_apply with nrhs = k in one
call; the result must equal k single-vector applies up to summation order
(FP64: 1e-13 relative per vector; FP32 copy: 2e-6 against its own vector path
and the FP64 result within the FP32 bound of the mixed matvec tests).
"""
import ctypes

import numpy as np
import pytest
import torch

import ngbem_amd as nb
from ngbem_amd import _capi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _ops():
    m3 = nb.icosphere(3)
    p0, p1, hd = nb.SurfaceL2(m3), nb.H1(m3), nb.HDivSurface(m3)
    kw = dict(intorder=3, leafsize=32, eta=2.0, eps=1e-6, mixed=True)
    return {
        "LSL": (nb.SingleLayerPotentialOperator(p0, **kw), O.LAPLACE_SL, 0.0, O.P0, O.P0),
        "LDL-P1P0": (nb.DoubleLayerPotentialOperator(p1, p0, **kw), O.LAPLACE_DL, 0.0, O.P1, O.P0),
        "HCF": (nb.HelmholtzCombinedFieldOperator(p0, p0, kappa=6.0, **kw), O.HELMHOLTZ_CF, 6.0, O.P0, O.P0),
        "HSL-P1": (nb.HelmholtzSingleLayerPotentialOperator(p1, p1, kappa=3.0, **kw), O.HELMHOLTZ_SL, 3.0, O.P1,
                   O.P1),
        "MSL": (nb.MaxwellSingleLayerPotentialOperator(hd, 4.0, **kw), O.MAXWELL_SL, 4.0, O.HDIV, O.HDIV),
    }


@pytest.fixture(scope="module")
def ops(gpu):
    return _ops()


@pytest.mark.parametrize("name", ["LSL", "LDL-P1P0", "HCF", "HSL-P1", "MSL"])
@pytest.mark.parametrize("k", [70, 5])
def test_block_equals_vector_path(gpu, ops, name, k):
    op, *_ = ops[name]
    print(f"block apply {name} k={k}", flush=True)
    dt = torch.complex128 if op.is_complex else torch.float64
    g = torch.Generator(device="cpu").manual_seed(3)
    X = torch.randn(k, op.trial_space.ndof, dtype=dt, generator=g).to(gpu)
    for A in (op.mat, op.mat.mixed):
        Y = A * X
        assert Y.shape == (k, op.test_space.ndof)
        Yv = torch.stack([A * X[r] for r in range(k)])
        err = (torch.linalg.norm(Y - Yv, dim=1) / torch.linalg.norm(Yv, dim=1)).max().item()
        assert err <= (1e-13 if A is op.mat else 2e-6), err
    Y64, Y32 = op.mat * X, op.mat.mixed * X
    assert (torch.linalg.norm(Y64 - Y32) / torch.linalg.norm(Y64)).item() < 1e-5


@pytest.mark.parametrize("name", ["LDL-P1P0", "HCF", "MSL"])
def test_block_matches_oracle(gpu, ops, name):
    op, kid, kappa, trk, tek = ops[name]
    ref = O.OracleOp(op.trial_space.mesh, kid, kappa=kappa, trial=trk, test=tek, intorder=3, leafsize=32, eta=2.0,
                     eps=1e-6)
    for b in range(op.num_blocks()):
        d = op.block_data(b)
        if isinstance(d, tuple):
            ref.set_block(b, np.concatenate([d[0].ravel(), d[1].ravel()]), d[0].shape[1])
        else:
            ref.set_block(b, d.ravel(), -1)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((3, op.trial_space.ndof)) + 1j * rng.standard_normal((3, op.trial_space.ndof))
    if not op.is_complex:
        X = X.real
    Y = (op.mat * torch.from_numpy(X).to(gpu)).cpu().numpy()
    for r in range(3):
        y_ref = ref.apply(X[r].astype(np.complex128))
        assert np.linalg.norm(Y[r] - y_ref) <= 1e-13 * np.linalg.norm(y_ref)


def test_block_apply_host_buffers_and_flags(gpu, ops):
    """Host pointers with ld > n, beta != 0, and the per-vector flag."""
    op, *_ = ops["HCF"]
    L = _capi.lib()
    n, k, ld = op.trial_space.ndof, 66, op.trial_space.ndof + 3
    rng = np.random.default_rng(9)
    X = np.zeros((k, ld), dtype=np.complex128)
    X[:, :n] = rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))
    Y0 = np.zeros((k, ld), dtype=np.complex128)
    Y0[:, :n] = rng.standard_normal((k, n))
    alpha = np.array([0.5, -1.0])
    beta = np.array([2.0, 0.25])
    out = {}
    for flags in (0, _capi.APPLY_PER_VECTOR):
        Y = Y0.copy()
        _capi.check(L.bem_op_apply(op._h, flags, _capi.dptr(alpha), ctypes.c_void_p(X.ctypes.data),
                                   _capi.dptr(beta), ctypes.c_void_p(Y.ctypes.data), k, ld, _capi.C128, None))
        out[flags] = Y
        assert np.array_equal(Y[:, n:], Y0[:, n:])  # padding untouched
    A = out[0][:, :n]
    B = out[_capi.APPLY_PER_VECTOR][:, :n]
    assert np.linalg.norm(A - B) <= 1e-13 * np.linalg.norm(B)
    Xt = torch.from_numpy(np.ascontiguousarray(X[:, :n])).to(gpu)
    ref = (0.5 - 1j) * (op.mat * Xt).cpu().numpy() + (2.0 + 0.25j) * Y0[:, :n]
    assert np.linalg.norm(A - ref) <= 1e-13 * np.linalg.norm(ref)
