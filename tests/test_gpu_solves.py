"""End-to-end solves on the GPU against analytic answers and the oracle.

- Laplace capacitor (single layer, SurfaceL2 p0, CG to 1e-10): capacitance of
  the unit sphere 4 pi; at level 3 equal to the oracle's dense solve.
- Laplace double layer + hypersingular on H1 p1: second-kind Dirichlet
  problem and stabilised hypersingular problem, FP64 vs mixed with FP64
  iterative refinement (solutions agree to 1e-6).
- Helmholtz combined field (Burton-Miller right-hand side) against the
  acoustic Mie series for the sound-soft sphere.
"""
import math

import numpy as np
import pytest
import torch

import ngbem_amd as nb
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_capacitance_level4(gpu):
    """Config 1: icosphere L4, intorder 4, leafsize 40, eta 3, eps 1e-8, CG 1e-10."""
    mesh = nb.icosphere(4)
    fes = nb.SurfaceL2(mesh, order=0)
    V = nb.SingleLayerPotentialOperator(fes, intorder=4, leafsize=40, eta=3.0, eps=1e-8)
    area = torch.from_numpy(mesh.areas()).to(gpu)
    sigma = nb.CG(V.mat, area, tol=1e-10, maxsteps=500, printrates=False)
    cap = float(area @ sigma)
    assert abs(cap - 4 * math.pi) / (4 * math.pi) < 2e-3  # flat-triangle geometry error
    r = area - V.mat * sigma
    assert float(torch.linalg.norm(r) / torch.linalg.norm(area)) < 1e-9


def test_capacitance_matches_oracle_level3(gpu):
    mesh = nb.icosphere(3)
    fes = nb.SurfaceL2(mesh, order=0)
    V = nb.SingleLayerPotentialOperator(fes, intorder=4, leafsize=40, eta=3.0, eps=1e-8)
    area_np = mesh.areas()
    sigma = nb.CG(V.mat, torch.from_numpy(area_np).to(gpu), tol=1e-12, maxsteps=500, printrates=False)
    D = O.OracleOp(mesh, O.LAPLACE_SL, intorder=4).dense().real
    cap_ref = area_np @ np.linalg.solve(D, area_np)
    assert abs(float(sigma.cpu().numpy() @ area_np) - cap_ref) / cap_ref < 1e-7


def _h1_data(mesh, seed=0):
    x, y, z = mesh.xyz.T
    g = x * y * z + 1e-3 * np.random.default_rng(seed).standard_normal(mesh.nv)
    return g


def test_dl_hs_h1_mixed_refinement(gpu):
    """Config 2 form on icosphere L4: DL and HS on H1 p1, intorder 4, eps 1e-6."""
    mesh = nb.icosphere(4)
    h1 = nb.H1(mesh, order=1)
    kw = dict(intorder=4, leafsize=40, eta=2.0, eps=1e-6, mixed=True)
    K = nb.DoubleLayerPotentialOperator(h1, h1, **kw)
    W = nb.HypersingularOperator(h1, **kw)
    M = nb.MassMatrix(h1)
    g = torch.from_numpy(_h1_data(mesh)).to(gpu)
    # second kind: (1/2 M + K) u = M g  (interior Dirichlet data of a harmonic function)
    A64 = 0.5 * M + K.mat
    A32 = 0.5 * M + K.mat.mixed
    f = M * g
    u64 = nb.GMRes(A64, f, tol=1e-12, maxsteps=300, printrates=False)
    u32, hist = nb.RefinedSolve(A64, A32, f, tol=1e-12, inner_tol=1e-5)
    assert hist[-1] <= 1e-12
    assert float(torch.linalg.norm(u64 - u32) / torch.linalg.norm(u64)) < 1e-6
    # stabilised hypersingular: (W + (M1)(M1)^T) v = M g
    ones = torch.ones(h1.ndof, dtype=torch.float64, device=gpu)
    m1 = M * ones

    class Stab(nb.BaseMatrix):
        def __init__(self, Wm):
            self.Wm, self.height, self.width, self.is_complex = Wm, Wm.height, Wm.width, False

        def MultAdd(self, s, x, y):
            self.Wm.MultAdd(s, x, y)
            y.add_(s * (m1.to(x.dtype) @ x) * m1.to(y.dtype))

    v64 = nb.CG(Stab(W.mat), f, tol=1e-12, maxsteps=500, printrates=False)
    v32, hist = nb.RefinedSolve(Stab(W.mat), Stab(W.mat.mixed), f, tol=1e-12, inner_tol=1e-5)
    assert float(torch.linalg.norm(v64 - v32.real) / torch.linalg.norm(v64)) < 1e-6
    # W annihilates constants exactly (surface curl of a constant is zero)
    w1 = W.mat * ones
    # exact for the dense part; the ACA blocks (eps 1e-6) keep it only to their tolerance
    assert float(torch.linalg.norm(w1)) <= 1e-4 * float(torch.linalg.norm(W.mat * g))


def mie_dudn(kappa, cos_t):
    from scipy.special import eval_legendre, spherical_jn, spherical_yn

    s = np.zeros_like(cos_t, dtype=np.complex128)
    for n in range(int(kappa) + 40):
        h = spherical_jn(n, kappa) + 1j * spherical_yn(n, kappa)
        s += (1j ** n) * (2 * n + 1) * eval_legendre(n, cos_t) / h
    return -1j / kappa * s


def test_burton_miller_vs_mie(gpu):
    """Sound-soft sphere, incident e^{i kappa z}: (1/2 M + C) u = f, u -> du/dn of the total field."""
    import bench

    errs = []
    for lvl in (3, 4):
        mesh = nb.icosphere(lvl)
        fes = nb.SurfaceL2(mesh, order=0)
        kappa = 4.0
        C = nb.HelmholtzCombinedFieldOperator(fes, fes, kappa=kappa, intorder=4, leafsize=40, eps=1e-6)
        f = torch.from_numpy(bench.burton_miller_rhs(mesh, kappa)).to(gpu)
        area = torch.from_numpy(mesh.areas()).to(gpu)
        u = nb.GMRes(0.5 * nb.MassMatrix(fes) + C.mat, f, pre=lambda r: r / area, tol=1e-10, maxsteps=200,
                     printrates=False)
        cen = mesh.xyz[mesh.tri].mean(1)
        ref = mie_dudn(kappa, cen[:, 2] / np.linalg.norm(cen, axis=1))
        w = mesh.areas()
        un = u.cpu().numpy()
        errs.append(math.sqrt(np.sum(w * np.abs(un - ref) ** 2) / np.sum(w * np.abs(ref) ** 2)))
    assert errs[1] < 0.6 * errs[0]
    assert errs[1] < 2e-2
