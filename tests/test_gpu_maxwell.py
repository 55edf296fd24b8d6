"""Maxwell on the GPU: the manufactured dipole problem of the reference's
tests/test_Maxwell_Box.py on the unit icosphere, solved with the HIP
H-matrices (single layer by GMRES, double layer on the right-hand side).

The exact current is known, and the oracle's dense matrices give the same
discrete solution up to the ACA tolerance (condition-amplified).
"""
import numpy as np
import pytest
import torch

import ngbem_amd as nb
from oracle import oracle as O
from test_oracle_maxwell import DipoleProblem, dense_mass

pytestmark = pytest.mark.gpu


def test_maxwell_dipole_level3(gpu):
    kappa, lvl = 3.0, 3
    mesh = nb.icosphere(lvl)
    pb = DipoleProblem(mesh, kappa)
    kw = dict(intorder=4, leafsize=40, eta=2.0, eps=1e-8)
    print(f"dipole L{lvl} kappa {kappa} ndof {pb.hd.ndof}", flush=True)
    V = nb.MaxwellSingleLayerPotentialOperator(pb.hd, kappa, **kw)
    K = nb.MaxwellDoubleLayerPotentialOperator(pb.hc, pb.hd, kappa, **kw)
    assert V.stats()["n_far"] > 0 and K.stats()["n_far"] > 0
    rhs = pb.rhs(lambda m: (K.mat * torch.from_numpy(m.astype(np.complex128)).to(gpu)).cpu().numpy())
    Minv = torch.linalg.inv(torch.from_numpy(dense_mass(pb.hd, pb.hd)).to(gpu)).to(torch.complex128)
    j = nb.GMRes(V.mat, torch.from_numpy(rhs).to(gpu), pre=lambda r: Minv @ r, tol=1e-11, maxsteps=2000,
                 printrates=False).cpu().numpy()
    err = pb.error(j)
    assert err < 6e-3  # second-order convergence: 0.059 (L1), 0.016 (L2), 0.0045 (L3)
    Vd = O.OracleOp(mesh, O.MAXWELL_SL, kappa, O.HDIV, O.HDIV, intorder=4).dense()
    Kd = O.OracleOp(mesh, O.MAXWELL_DL, kappa, O.HCURL, O.HDIV, intorder=4).dense()
    jd = np.linalg.solve(Vd, pb.rhs(lambda m: Kd @ m))
    assert np.linalg.norm(j - jd) <= 1e-4 * np.linalg.norm(jd)
    assert abs(err - pb.error(jd)) <= 1e-3 * err


def test_maxwell_mixed_refinement(gpu):
    """FP32 H-matrix with FP64 refinement reaches the FP64 solution."""
    kappa = 3.0
    mesh = nb.icosphere(3)
    hd = nb.HDivSurface(mesh)
    V = nb.MaxwellSingleLayerPotentialOperator(hd, kappa, intorder=3, leafsize=40, eps=1e-6, mixed=True)
    Minv = torch.linalg.inv(torch.from_numpy(dense_mass(hd, hd)).to(gpu)).to(torch.complex128)
    pre = lambda r: Minv @ r  # noqa: E731
    b = torch.randn(hd.ndof, dtype=torch.complex128, device=gpu)
    x64 = nb.GMRes(V.mat, b, pre=pre, tol=1e-11, maxsteps=2000, printrates=False)
    x32, hist = nb.RefinedSolve(V.mat, V.mat.mixed, b, pre=pre, tol=1e-11, inner_tol=1e-4, maxsteps=2000)
    assert hist[-1] <= 1e-11
    assert float(torch.linalg.norm(x64 - x32) / torch.linalg.norm(x64)) < 1e-6


def plane_waves(nrhs, seed=0):
    """Directions uniform on the sphere, polarisations random and orthogonal."""
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((nrhs, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    p = rng.standard_normal((nrhs, 3))
    p -= (p * d).sum(1)[:, None] * d
    p /= np.linalg.norm(p, axis=1)[:, None]
    return d, p


def test_efie_64_plane_waves_vs_mie(gpu):
    """Maxwell EFIE, kappa 5, HDivSurface p0, intorder 4, ACA eps 1e-5, 64
    plane waves solved together (batched GMRES: every step is one block
    product on the matrix cores); each current against its Mie series."""
    from test_oracle_maxwell import efie_current_error, efie_plane_wave_rhs

    k, lvl = 5.0, 4
    mesh = nb.icosphere(lvl)
    hd = nb.HDivSurface(mesh)
    print(f"EFIE L{lvl} ndof {hd.ndof} 64 plane waves", flush=True)
    V = nb.MaxwellSingleLayerPotentialOperator(hd, k, intorder=4, leafsize=40, eta=2.0, eps=1e-5)
    d, p = plane_waves(64)
    B = torch.from_numpy(efie_plane_wave_rhs(hd, k, d, p)).to(gpu)
    MinvT = torch.linalg.inv(torch.from_numpy(dense_mass(hd, hd)).to(gpu)).T.contiguous()
    pre = lambda R: torch.complex(R.real @ MinvT, R.imag @ MinvT)  # noqa: E731
    J = nb.GMResBatch(V.mat, B, pre=pre, tol=1e-6, maxsteps=1500, restart=300)
    print(f"batched GMRES iterations {nb.GMResBatch.iterations}", flush=True)
    assert np.all(nb.GMResBatch.residuals <= 1e-6)
    errs = efie_current_error(hd, J.cpu().numpy(), k, d, p)
    print(f"current errors: mean {errs.mean():.4f} max {errs.max():.4f}", flush=True)
    # the oracle's dense solve at L4 gives 0.074 for the wave along +z (0.149 at L3: first order)
    assert errs.max() < 0.09 and errs.mean() < 0.08
