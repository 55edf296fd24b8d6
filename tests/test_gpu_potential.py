"""GetPotential (PotentialCF) on the GPU against the oracle's restatement and
Green's representation formula."""
import numpy as np
import pytest
import torch

import ngbem_amd as nb
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    ("LSL", nb.SingleLayerPotentialOperator, O.LAPLACE_SL, 0, 0, 0.0),
    ("LDL", nb.DoubleLayerPotentialOperator, O.LAPLACE_DL, 1, 0, 0.0),
    ("LHS", nb.HypersingularOperator, O.LAPLACE_HS, 1, 1, 0.0),
    ("HSL", nb.HelmholtzSingleLayerPotentialOperator, O.HELMHOLTZ_SL, 0, 0, 3.0),
    ("HCF", nb.HelmholtzCombinedFieldOperator, O.HELMHOLTZ_CF, 0, 0, 3.0),
    ("HHS", nb.HelmholtzHypersingularOperator, O.HELMHOLTZ_HS, 1, 1, 3.0),
    ("MSL", nb.MaxwellSingleLayerPotentialOperator, O.MAXWELL_SL, 2, 2, 3.0),
    ("MDL", nb.MaxwellDoubleLayerPotentialOperator, O.MAXWELL_DL, 3, 2, 3.0),
]


def _space(mesh, kind):
    return [nb.SurfaceL2, nb.H1, nb.HDivSurface, nb.HCurl][kind](mesh)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_potential_matches_oracle(gpu, case):
    name, fac, kid, trk, tek, kappa = case
    mesh = nb.icosphere(2)
    trial, test = _space(mesh, trk), _space(mesh, tek)
    kw = dict(intorder=4, leafsize=32, eta=2.0, eps=1e-8)
    print(f"potential {name}", flush=True)
    if name in ("LSL", "LHS"):
        op = fac(trial, **kw)
    elif name == "LDL":
        op = fac(trial, test, **kw)
    elif name == "MSL":
        op = fac(trial, kappa, **kw)
    elif name == "MDL":
        op = fac(trial, test, kappa, **kw)
    else:
        op = fac(trial, test, kappa=kappa, **kw)
    rng = np.random.default_rng(4)
    coef = rng.standard_normal(trial.ndof) + (1j * rng.standard_normal(trial.ndof) if op.is_complex else 0)
    pts = rng.standard_normal((300, 3))
    pts *= (0.2 + 2.0 * rng.random(300))[:, None] / np.linalg.norm(pts, axis=1)[:, None]
    pts = pts[np.abs(np.linalg.norm(pts, axis=1) - 1.0) > 0.05]
    nrm = rng.standard_normal(pts.shape)
    pot = op.GetPotential(torch.from_numpy(coef).to(gpu))
    ref = O.OracleOp(mesh, kid, kappa=kappa, trial=trk, test=tek, **kw)
    for normals in (None, nrm):
        u = pot(torch.from_numpy(pts).to(gpu), None if normals is None else torch.from_numpy(normals).to(gpu))
        u = u.cpu().numpy().reshape(len(pts), -1)
        r = ref.potential(coef, pts, normals)
        assert u.shape == r.shape == (len(pts), pot.dim)
        if not op.is_complex:
            assert np.abs(r.imag).max() == 0.0
            r = r.real
        assert np.abs(u - r).max() <= 1e-12 * np.abs(r).max()


def test_representation_formula_level4(gpu):
    """u = SL(du/dn) - DL(u) inside for the field of an outside point source."""
    x0 = np.array([0.3, -0.2, 2.5])
    mesh = nb.icosphere(4)
    fes = nb.SurfaceL2(mesh)
    rule = O.trig_rule(6)
    bary = np.stack([rule[:, 0], rule[:, 1], 1 - rule[:, 0] - rule[:, 1]], 1)
    X = np.einsum("qk,tkc->tqc", bary, mesh.xyz[mesh.tri])
    w = rule[:, 2] * 2
    D = X - x0
    R = np.linalg.norm(D, axis=-1)
    u = (w[None] / (4 * np.pi * R)).sum(1) / w.sum()
    dn = (w[None] * (-(D * mesh.normals()[:, None, :]).sum(-1) / (4 * np.pi * R ** 3))).sum(1) / w.sum()
    V = nb.SingleLayerPotentialOperator(fes, intorder=4)
    K = nb.DoubleLayerPotentialOperator(fes, fes, intorder=4)
    ip = np.array([[0.1, 0.2, -0.1], [0.3, -0.4, 0.2], [-0.5, 0.1, 0.1]])
    P = torch.from_numpy(ip).to(gpu)
    rep = V.GetPotential(torch.from_numpy(dn).to(gpu))(P) - K.GetPotential(torch.from_numpy(u).to(gpu))(P)
    ex = 1.0 / (4 * np.pi * np.linalg.norm(ip - x0, axis=1))
    assert np.abs(rep.cpu().numpy() - ex).max() / ex.max() < 5e-5  # 2.5e-5 measured with the oracle at L4
