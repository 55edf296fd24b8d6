"""Batches of operators (one launch sequence for many geometries, replayed
from a hipGraph) and the differentiable single layer (adjoint apply and the
d/dkappa operator behind torch.autograd), on the GPU.

Bars: batch apply = per-instance applies within 1e-13 (same kernels, same
summation order per row); graph replays = eager launches exactly, replay after
replay with other GPU work in between; gradients against central finite
differences with h = 1e-6 in FP64 (relative 1e-6), on operators without
low-rank blocks so that ACA does not enter the difference quotient.
"""
import numpy as np
import pytest
import torch

import ngbem_amd as nb
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _instances(n, levels=(2, 3), seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        axes = rng.uniform(0.5, 2.0, 3)
        lvl = int(rng.choice(levels))
        kappa = float(rng.uniform(1.0, 8.0))
        out.append((nb.ellipsoid(lvl, tuple(axes)), kappa))
    return out


@pytest.fixture(scope="module")
def batch_setup(gpu):
    inst = _instances(6)
    spaces = [nb.SurfaceL2(m) for m, _ in inst]
    kw = dict(intorder=3, leafsize=32, eta=2.0, eps=1e-8, mixed=True)
    ops = nb.assemble_many(lambda it: nb.HelmholtzSingleLayerPotentialOperator(it[0], it[0], kappa=it[1], **kw),
                           list(zip(spaces, [k for _, k in inst])), threads=4)
    return inst, spaces, ops, nb.OperatorBatch(ops)


def test_batch_equals_instances(gpu, batch_setup):
    inst, spaces, ops, B = batch_setup
    print(f"batch of {len(ops)}: ndofs {[s.ndof for s in spaces]}", flush=True)
    assert B.width == sum(s.ndof for s in spaces)
    x = torch.randn(B.width, dtype=torch.complex128, device=gpu)
    for graph in (True, False):
        for mixed in (False, True):
            V = B.view(mixed=mixed, graph=graph)
            for kind in ("fwd", "trans", "adj"):
                M = V if kind == "fwd" else (V.T if kind == "trans" else V.H)
                y = M * x
                parts = []
                for op, xi in zip(ops, B.split(x, rows=False)):
                    A = op.mat.mixed if mixed else op.mat
                    A = A if kind == "fwd" else (A.T if kind == "trans" else A.H)
                    parts.append(A * xi)
                ref = torch.cat(parts)
                tol = 1e-13 if not mixed else 1e-6
                assert (torch.linalg.norm(y - ref) / torch.linalg.norm(ref)).item() <= tol, (graph, mixed, kind)


def test_graph_replays_match_eager(gpu, batch_setup):
    _, _, ops, B = batch_setup
    eager = B.view(graph=False)
    g = torch.Generator(device="cpu").manual_seed(7)
    for rep in range(6):
        x = torch.randn(B.width, dtype=torch.complex128, generator=g).to(gpu)
        y_ref = eager * x
        torch.cuda.synchronize()
        _ = torch.randn(4096, 4096, device=gpu).sum()  # other work between replays
        y = B * x
        y2 = torch.zeros_like(y)
        B.MultAdd(2.0 - 1.0j, x, y2)
        torch.cuda.synchronize()
        assert torch.equal(y, y_ref), rep
        assert (torch.linalg.norm(y2 - (2.0 - 1.0j) * y_ref) / torch.linalg.norm(y_ref)).item() < 1e-15


def test_dkappa_operator_matches_oracle(gpu):
    m = nb.ellipsoid(2, (1.2, 0.8, 0.6))
    p0 = nb.SurfaceL2(m)
    kw = dict(intorder=4, leafsize=16, eta=2.0, eps=1e-10)
    D = nb.HelmholtzSingleLayerPotentialOperatorDKappa(p0, p0, kappa=3.5, **kw)
    ref = O.OracleOp(m, O.HELMHOLTZ_SL_DKAPPA, 3.5, O.P0, O.P0, **kw)
    Dd = ref.dense()
    x = np.random.default_rng(3).standard_normal(p0.ndof) + 0j
    y = (D.mat * torch.from_numpy(x).to(gpu)).cpu().numpy()
    assert np.linalg.norm(y - Dd @ x) <= 1e-8 * np.linalg.norm(Dd @ x)
    for b in range(D.num_blocks()):
        if D.block_info(b)["rank"] < 0:
            rows, cols = D.block_dofs(b)
            assert np.abs(D.block_data(b) - ref.dense(cols, rows)).max() <= 1e-12 * np.abs(Dd).max()


def _loss(slb, x, kappa, target):
    y = slb(x, kappa)
    return ((y - target).abs() ** 2).sum()


def test_autograd_against_finite_differences(gpu):
    """loss = ||V(kappa) x - t||^2 over a batch; d loss / d kappa_i and the
    directional derivative in x against central differences (h = 1e-6)."""
    inst = _instances(4, levels=(1, 2), seed=3)
    spaces = [nb.SurfaceL2(m) for m, _ in inst]
    kap = [k for _, k in inst]
    kw = dict(intorder=4, leafsize=100000, eta=2.0, eps=1e-10)  # all blocks dense: no ACA in the difference
    slb = nb.SingleLayerBatch(spaces, kap, threads=2, **kw)
    n = sum(s.ndof for s in spaces)
    g = torch.Generator(device="cpu").manual_seed(11)
    x = torch.randn(n, dtype=torch.float64, generator=g).to(gpu).requires_grad_(True)
    t = torch.randn(n, dtype=torch.complex128, generator=g).to(gpu)
    kappa = torch.tensor(kap, dtype=torch.float64, device=gpu, requires_grad=True)
    L = _loss(slb, x, kappa, t)
    L.backward()
    h = 1e-6
    for i in range(len(kap)):
        kp, km = list(kap), list(kap)
        kp[i] += h
        km[i] -= h
        Lp = _loss(nb.SingleLayerBatch(spaces, kp, threads=2, **kw), x.detach(), torch.tensor(kp, dtype=torch.float64), t).item()
        Lm = _loss(nb.SingleLayerBatch(spaces, km, threads=2, **kw), x.detach(), torch.tensor(km, dtype=torch.float64), t).item()
        fd = (Lp - Lm) / (2 * h)
        assert abs(kappa.grad[i].item() - fd) <= 1e-6 * max(abs(fd), 1e-3 * L.item()), (i, kappa.grad[i].item(), fd)
    d = torch.randn(n, dtype=torch.float64, generator=g).to(gpu)
    with torch.no_grad():
        Lp = _loss(slb, x + h * d, kappa, t).item()
        Lm = _loss(slb, x - h * d, kappa, t).item()
    fd = (Lp - Lm) / (2 * h)
    ad = torch.dot(x.grad, d).item()
    assert abs(ad - fd) <= 1e-6 * abs(fd)


def test_autograd_hmatrix_gradient_close_to_dense(gpu):
    """With low-rank blocks the kappa-gradient differs from the dense one only
    by the ACA tolerance."""
    inst = _instances(3, levels=(3,), seed=5)
    spaces = [nb.SurfaceL2(m) for m, _ in inst]
    kap = [k for _, k in inst]
    n = sum(s.ndof for s in spaces)
    g = torch.Generator(device="cpu").manual_seed(2)
    x0 = torch.randn(n, dtype=torch.float64, generator=g).to(gpu)
    t = torch.randn(n, dtype=torch.complex128, generator=g).to(gpu)
    grads = []
    for kw in (dict(intorder=3, leafsize=32, eta=2.0, eps=1e-8), dict(intorder=3, leafsize=100000, eta=2.0, eps=1e-8)):
        slb = nb.SingleLayerBatch(spaces, kap, threads=3, **kw)
        x = x0.clone().requires_grad_(True)
        kappa = torch.tensor(kap, dtype=torch.float64, device=gpu, requires_grad=True)
        _loss(slb, x, kappa, t).backward()
        grads.append((kappa.grad.clone(), x.grad.clone()))
    assert (torch.linalg.norm(grads[0][0] - grads[1][0]) / torch.linalg.norm(grads[1][0])).item() < 1e-5
    assert (torch.linalg.norm(grads[0][1] - grads[1][1]) / torch.linalg.norm(grads[1][1])).item() < 1e-5


def test_consumed_batch_equals_copied_batch(gpu, batch_setup):
    inst, spaces, ops, B = batch_setup
    kw = dict(intorder=3, leafsize=32, eta=2.0, eps=1e-8)
    fresh = [nb.HelmholtzSingleLayerPotentialOperator(s, s, kappa=k, **kw) for s, (_, k) in zip(spaces, inst)]
    C = nb.OperatorBatch(fresh, consume=True)
    x = torch.randn(B.width, dtype=torch.complex128, device=gpu)
    assert torch.equal(C * x, B * x)
    with pytest.raises(RuntimeError, match="not assembled"):
        fresh[0].mat * x[: spaces[0].ndof]


@pytest.mark.parametrize("kernel,trial,test", [
    ("HSL", nb.SurfaceL2, None), ("HHS", nb.H1, None), ("LDL", nb.H1, nb.SurfaceL2), ("HCF", nb.SurfaceL2, None)])
def test_multi_geometry_operator_equals_single_operators(gpu, kernel, trial, test):
    """One operator over several geometries (one assembly, one apply) gives
    every geometry the result of its own operator, with its own wavenumber."""
    from ngbem_amd import _capi

    inst = _instances(5, levels=(2, 3), seed=9)
    meshes = [m for m, _ in inst]
    kap = [k for _, k in inst]
    kid = {"HSL": _capi.HELMHOLTZ_SL, "HHS": _capi.HELMHOLTZ_HS, "LDL": _capi.LAPLACE_DL,
           "HCF": _capi.HELMHOLTZ_CF}[kernel]
    fac = {"HSL": nb.HelmholtzSingleLayerPotentialOperator, "HHS": nb.HelmholtzHypersingularOperator,
           "LDL": nb.DoubleLayerPotentialOperator, "HCF": nb.HelmholtzCombinedFieldOperator}[kernel]
    kw = dict(intorder=3, leafsize=24, eta=2.0, eps=1e-8)
    print(f"multi-geometry {kernel}: {[m.nt for m in meshes]} triangles", flush=True)
    M = nb.MultiGeometryOperator(kid, meshes, kap if kid >= _capi.HELMHOLTZ_SL else None, trial, test, **kw)
    singles = []
    for m, k in zip(meshes, kap):
        tr, te = trial(m), (test(m) if test else None)
        if kernel == "LDL":
            singles.append(fac(tr, te, **kw))
        elif kernel == "HHS":
            singles.append(fac(tr, kappa=k, **kw))
        else:
            singles.append(fac(tr, tr, kappa=k, **kw))
    toff = np.cumsum([0] + [s.test_space.ndof for s in singles])
    coff = np.cumsum([0] + [s.trial_space.ndof for s in singles])
    dt = torch.complex128 if M.is_complex else torch.float64
    g = torch.Generator(device="cpu").manual_seed(1)
    x = torch.randn(M.trial_space.ndof, dtype=dt, generator=g).to(gpu)
    xt = torch.randn(M.test_space.ndof, dtype=dt, generator=g).to(gpu)
    y = M.mat * x
    yt = M.mat.T * xt
    ref = torch.cat([s.mat * x[coff[i]:coff[i + 1]] for i, s in enumerate(singles)])
    ref_t = torch.cat([s.mat.T * xt[toff[i]:toff[i + 1]] for i, s in enumerate(singles)])
    assert (torch.linalg.norm(y - ref) / torch.linalg.norm(ref)).item() <= 1e-13
    assert (torch.linalg.norm(yt - ref_t) / torch.linalg.norm(ref_t)).item() <= 1e-13
    st = M.stats()
    assert st["n_near"] == sum(s.stats()["n_near"] for s in singles)
    assert st["n_far"] == sum(s.stats()["n_far"] for s in singles)
