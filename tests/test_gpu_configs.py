"""The benchmark configurations at their stated sizes, on the GPU.

- Laplace double layer + hypersingular, H1 order 1 on the icosphere of level 5
  (20,480 triangles, 10,242 dofs), intorder 4, leafsize 40, eta 2, ACA eps
  1e-6, data g = xyz + 1e-3 N(0, 1) (seed 0): a seeded 0.5% of the blocks
  against the oracle, FP64 against mixed + FP64 refinement (1e-6).
- Helmholtz combined field, the headline: icosphere level 6 (81,920
  triangles), kappa 10, SurfaceL2 0, intorder 4, leafsize 64, eta 2, eps 1e-6:
  a seeded 0.5% of the blocks against the oracle (dense 1e-12, low rank 10
  eps), the apply against the oracle's apply on the GPU's blocks (1e-13), the
  mixed and FP64 Burton-Miller solutions against each other (1e-6) and the
  FP64 one against the acoustic Mie series (1.5e-3).
- Maxwell EFIE, kappa 5, HDivSurface 0 on the icosphere of level 5 (30,720
  dofs), intorder 4, eps 1e-5: 64 plane waves (seed 0) solved together through
  the block product on the matrix cores, every current against its Mie series.
- 256 Helmholtz single layers on ellipsoids (semi-axes U[0.5, 2], levels
  U{3, 4, 5}, kappa U[1, 8], seed 0) in one multi-geometry operator: sampled
  instances against their own operators (1e-13), and the kappa- and
  x-gradients of ||V(kappa) x - t||^2 of sampled instances against central
  differences (h = 1e-6, FP64) of the dense single-instance operators: the
  H-matrix gradient differs from the dense one by the ACA tolerance times
  the gradient's sensitivity scale 2 |V x - t| |dV x| (10 eps).
"""
import math
import time

import numpy as np
import pytest
import torch

import ngbem_amd as nb
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _release_cached_memory():
    """These workloads use tens of GB each: give memory that torch's caching
    allocator keeps back to the device between tests, so the library's own
    allocations see it."""
    yield
    import gc

    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def sampled_block_parity(op, ref, eps, frac=0.005, seed=0):
    """A seeded fraction of the blocks (at least 20 of each kind): dense blocks
    within 1e-12 of the oracle's entries (relative to the block's largest),
    low-rank blocks within 10 eps of the oracle's cross approximation; returns
    (dense checked, low rank checked, same rank)."""
    rng = np.random.default_rng(seed)
    infos = ref.blocks()
    near = [b for b, i in enumerate(infos) if i["near"] or i["fallback"]]
    far = [b for b, i in enumerate(infos) if not i["near"] and not i["fallback"]]
    pick = lambda ids: rng.choice(ids, size=min(len(ids), max(20, int(frac * len(ids)))), replace=False)  # noqa: E731
    nd = nf = same = 0
    for b in pick(near):
        rows, cols = op.block_dofs(b)
        D = ref.dense(cols, rows)
        data = op.block_data(b)
        if not op.is_complex:
            D = D.real
        assert np.abs(data - D).max() <= 1e-12 * np.abs(D).max(), f"dense block {b}"
        nd += 1
    for b in (pick(far) if far else []):
        rows, cols = op.block_dofs(b)
        U, V = ref.aca(cols, rows)
        d = op.block_data(b)
        A_ref = U @ V
        assert np.linalg.norm(d[0] @ d[1] - A_ref) <= 10 * eps * np.linalg.norm(A_ref), f"low-rank block {b}"
        same += d[0].shape[1] == U.shape[1]
        nf += 1
    assert same >= 0.9 * nf
    print(f"sampled blocks: {nd} dense, {nf} low rank ({same} at the oracle's rank)", flush=True)
    return nd, nf, same


def test_laplace_dl_hs_level5(gpu):
    mesh = nb.icosphere(5)
    h1 = nb.H1(mesh, order=1)
    assert h1.ndof == 10242
    kw = dict(intorder=4, leafsize=40, eta=2.0, eps=1e-6)
    print(f"Laplace DL + HS, icosphere L5, H1 1, {h1.ndof} dofs", flush=True)
    K = nb.DoubleLayerPotentialOperator(h1, h1, mixed=True, **kw)
    W = nb.HypersingularOperator(h1, mixed=True, **kw)
    sampled_block_parity(K, O.OracleOp(mesh, O.LAPLACE_DL, trial=O.P1, test=O.P1, **kw), kw["eps"])
    sampled_block_parity(W, O.OracleOp(mesh, O.LAPLACE_HS, trial=O.P1, test=O.P1, **kw), kw["eps"], seed=1)
    M = nb.MassMatrix(h1)
    x, y, z = mesh.xyz.T
    g = torch.from_numpy(x * y * z + 1e-3 * np.random.default_rng(0).standard_normal(mesh.nv)).to(gpu)
    f = M * g
    A64, A32 = 0.5 * M + K.mat, 0.5 * M + K.mat.mixed
    u64 = nb.GMRes(A64, f, tol=1e-12, maxsteps=400, printrates=False)
    u32, hist = nb.RefinedSolve(A64, A32, f, tol=1e-12, inner_tol=1e-5)
    assert hist[-1] <= 1e-12
    du = float(torch.linalg.norm(u64 - u32) / torch.linalg.norm(u64))
    ones = torch.ones(h1.ndof, dtype=torch.float64, device=gpu)
    m1 = M * ones

    class Stab(nb.BaseMatrix):  # W + (M 1)(M 1)^T: the hypersingular operator made regular
        def __init__(self, Wm):
            self.Wm, self.height, self.width, self.is_complex = Wm, Wm.height, Wm.width, False

        def MultAdd(self, s, xx, yy):
            self.Wm.MultAdd(s, xx, yy)
            yy.add_(s * (m1.to(xx.dtype) @ xx) * m1.to(yy.dtype))

    v64 = nb.CG(Stab(W.mat), f, tol=1e-12, maxsteps=1000, printrates=False)
    v32, hist = nb.RefinedSolve(Stab(W.mat), Stab(W.mat.mixed), f, tol=1e-12, inner_tol=1e-5)
    dv = float(torch.linalg.norm(v64 - v32.real) / torch.linalg.norm(v64))
    print(f"mixed + refinement vs FP64: DL solution {du:.1e}, HS solution {dv:.1e}", flush=True)
    assert du < 1e-6 and dv < 1e-6


def test_helmholtz_cf_headline_level6(gpu):
    import bench

    mesh = nb.icosphere(6)
    fes = nb.SurfaceL2(mesh)
    kappa = 10.0
    kw = dict(intorder=4, leafsize=64, eta=2.0, eps=1e-6)
    print(f"Helmholtz combined field, icosphere L6, {fes.ndof} dofs, kappa {kappa}", flush=True)
    C = nb.HelmholtzCombinedFieldOperator(fes, fes, kappa=kappa, mixed=True, **kw)
    ref = O.OracleOp(mesh, O.HELMHOLTZ_CF, kappa=kappa, trial=O.P0, test=O.P0, **kw)
    assert C.num_blocks() == ref.nblocks
    sampled_block_parity(C, ref, kw["eps"])
    # the apply against the oracle's apply on the GPU's blocks
    for b in range(C.num_blocks()):
        d = C.block_data(b)
        if isinstance(d, tuple):
            ref.set_block(b, np.concatenate([d[0].ravel(), d[1].ravel()]), d[0].shape[1])
        else:
            ref.set_block(b, d.ravel(), -1)
    x = np.random.default_rng(1).standard_normal(fes.ndof) + 1j * np.random.default_rng(2).standard_normal(fes.ndof)
    y = (C.mat * torch.from_numpy(x).to(gpu)).cpu().numpy()
    y_ref = ref.apply(x, nthreads=16)
    err_apply = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    yt = (C.mat.T * torch.from_numpy(x).to(gpu)).cpu().numpy()
    yt_ref = ref.apply(x, trans=True, nthreads=16)
    err_t = np.linalg.norm(yt - yt_ref) / np.linalg.norm(yt_ref)
    print(f"apply vs oracle apply on the same blocks: {err_apply:.1e} (transpose {err_t:.1e})", flush=True)
    assert err_apply <= 1e-13 and err_t <= 1e-13
    del ref
    # Burton-Miller solve, both forms to GMRES tolerance 1e-7
    f = torch.from_numpy(bench.burton_miller_rhs(mesh, kappa)).to(gpu)
    area = torch.from_numpy(mesh.areas()).to(gpu)
    Mass = nb.MassMatrix(fes)
    pre = lambda r: r / area  # noqa: E731
    u64 = nb.GMRes(0.5 * Mass + C.mat, f, pre=pre, tol=1e-7, maxsteps=200, printrates=False)
    u32 = nb.GMRes(0.5 * Mass + C.mat.mixed, f, pre=pre, tol=1e-7, maxsteps=200, printrates=False)
    diff = float(torch.linalg.norm(u64 - u32) / torch.linalg.norm(u64))
    cen = mesh.xyz[mesh.tri].mean(1)
    mie = bench.mie_dudn_soft_sphere(kappa, cen[:, 2] / np.linalg.norm(cen, axis=1))
    w = mesh.areas()
    un = u64.cpu().numpy()
    mie_err = math.sqrt(np.sum(w * np.abs(un - mie) ** 2) / np.sum(w * np.abs(mie) ** 2))
    print(f"mixed vs FP64 solution {diff:.1e}; FP64 against the Mie series {mie_err:.3e}", flush=True)
    assert diff <= 1e-6
    assert mie_err <= 1.5e-3


def test_maxwell_efie_level5_64_plane_waves(gpu):
    from test_gpu_maxwell import plane_waves
    from test_oracle_maxwell import dense_mass, efie_current_error, efie_plane_wave_rhs

    k = 5.0
    mesh = nb.icosphere(5)
    hd = nb.HDivSurface(mesh)
    assert hd.ndof == 30720
    print(f"EFIE L5 ndof {hd.ndof} 64 plane waves", flush=True)
    V = nb.MaxwellSingleLayerPotentialOperator(hd, k, intorder=4, leafsize=40, eta=2.0, eps=1e-5)
    d, p = plane_waves(64)
    B = torch.from_numpy(efie_plane_wave_rhs(hd, k, d, p)).to(gpu)
    MinvT = torch.linalg.inv(torch.from_numpy(dense_mass(hd, hd)).to(gpu)).T.contiguous()
    pre = lambda R: torch.complex(R.real @ MinvT, R.imag @ MinvT)  # noqa: E731
    t0 = time.perf_counter()
    J = nb.GMResBatch(V.mat, B, pre=pre, tol=1e-6, maxsteps=3000, restart=200)
    print(f"batched GMRES iterations {nb.GMResBatch.iterations} ({time.perf_counter() - t0:.1f} s)", flush=True)
    del MinvT, pre
    t0 = time.perf_counter()
    assert np.all(nb.GMResBatch.residuals <= 1e-6)
    errs = efie_current_error(hd, J.cpu().numpy(), k, d, p, order=4, progress=True)
    print(f"current errors against the Mie series: mean {errs.mean():.4f} max {errs.max():.4f} "
          f"({time.perf_counter() - t0:.1f} s)", flush=True)
    # first order in h: 0.149 (L3), 0.074 (L4) for the wave along +z
    assert errs.max() < 0.05 and errs.mean() < 0.045


def _ellipsoid_draws(n=256, seed=0):
    rng = np.random.default_rng(seed)
    axes = rng.uniform(0.5, 2.0, (n, 3))
    levels = rng.integers(3, 6, n)
    kappas = rng.uniform(1.0, 8.0, n)
    return axes, levels, kappas


def test_256_ellipsoids_batch_and_gradient(gpu):
    axes, levels, kappas = _ellipsoid_draws()
    meshes = [nb.ellipsoid(int(l), tuple(a)) for l, a in zip(levels, axes)]
    spaces = [nb.SurfaceL2(m) for m in meshes]
    kw = dict(intorder=4, leafsize=40, eta=2.0, eps=1e-6)
    print(f"256 ellipsoids: levels {np.bincount(levels, minlength=6)[3:].tolist()} (L3/L4/L5), "
          f"{sum(m.nt for m in meshes)} triangles", flush=True)
    slb = nb.SingleLayerBatch(spaces, kappas, **kw)
    off = np.cumsum([0] + [s.ndof for s in spaces])
    g = torch.Generator(device="cpu").manual_seed(3)
    n = int(off[-1])
    x = torch.randn(n, dtype=torch.complex128, generator=g).to(gpu)
    y = slb.V * x
    rng = np.random.default_rng(1)
    sample = [int(i) for i in rng.choice(256, size=4, replace=False)]
    sample += [int(np.nonzero(levels == 5)[0][0])]  # at least one of level 5
    for i in sample:
        Vi = nb.HelmholtzSingleLayerPotentialOperator(spaces[i], spaces[i], kappa=float(kappas[i]), **kw)
        yi = Vi.mat * x[off[i]:off[i + 1]]
        e = (torch.linalg.norm(y[off[i]:off[i + 1]] - yi) / torch.linalg.norm(yi)).item()
        assert e <= 1e-13, (i, e)
    # gradients of L = sum_i ||V_i(kappa_i) x_i - t_i||^2 for the whole batch
    xr = torch.randn(n, dtype=torch.float64, generator=g).to(gpu).requires_grad_(True)
    t = torch.randn(n, dtype=torch.complex128, generator=g).to(gpu)
    kap = torch.tensor(kappas, dtype=torch.float64, device=gpu, requires_grad=True)
    yb = slb(xr, kap)
    L = ((yb - t).abs() ** 2).sum()
    L.backward()
    res = (yb - t).detach()
    dVx = slb.dV * xr.detach().to(torch.complex128)
    # central differences of instance i alone, all blocks dense (no ACA in the
    # quotient); L's instance-i term is all that depends on kappa_i and x_i
    h = 1e-6
    dense = dict(kw, leafsize=10 ** 6)
    for i in [j for j in sample if levels[j] == 3][:2] + [int(np.nonzero(levels == 3)[0][0])]:
        xi, ti = xr.detach()[off[i]:off[i + 1]], t[off[i]:off[i + 1]]

        def Li(kappa, xv=xi):
            Vi = nb.HelmholtzSingleLayerPotentialOperator(spaces[i], spaces[i], kappa=kappa, **dense)
            return ((Vi.mat * xv.to(torch.complex128) - ti).abs() ** 2).sum().item()

        fd = (Li(float(kappas[i]) + h) - Li(float(kappas[i]) - h)) / (2 * h)
        ad = kap.grad[i].item()
        dvec = torch.randn(off[i + 1] - off[i], dtype=torch.float64, generator=g).to(gpu)
        fdx = (Li(float(kappas[i]), xi + h * dvec) - Li(float(kappas[i]), xi - h * dvec)) / (2 * h)
        adx = torch.dot(xr.grad[off[i]:off[i + 1]], dvec).item()
        # the gradient is 2 Re <V x - t, dV x> (and 2 Re <V x - t, V d>): an
        # operator error of relative size eps moves it by up to eps times
        # 2 |V x - t| |dV x| (2 |V x - t| |V d|), which bounds the H-matrix
        # gradient's distance from the dense one
        ri = res[off[i]:off[i + 1]]
        Vd = nb.HelmholtzSingleLayerPotentialOperator(spaces[i], spaces[i], kappa=float(kappas[i]), **dense).mat * \
            dvec.to(torch.complex128)
        sk = 2 * (torch.linalg.norm(ri) * torch.linalg.norm(dVx[off[i]:off[i + 1]])).item()
        sx = 2 * (torch.linalg.norm(ri) * torch.linalg.norm(Vd)).item()
        print(f"instance {i} (level {levels[i]}): d/dkappa {ad:.8e} vs FD {fd:.8e} (|diff| / scale "
              f"{abs(ad - fd) / sk:.1e}); d/dx.d {adx:.8e} vs FD {fdx:.8e} ({abs(adx - fdx) / sx:.1e})", flush=True)
        assert abs(ad - fd) <= 10 * kw["eps"] * sk, (i, ad, fd)
        assert abs(adx - fdx) <= 10 * kw["eps"] * sx, (i, adx, fdx)
