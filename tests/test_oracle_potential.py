"""Pins of the oracle's PotentialCF restatement (src/ngbem.cpp:1166-1235).

- Laplace single layer of the density 1: outside the closed surface the
  potential is the monopole |Gamma_h| / (4 pi |x|) up to the (tiny) higher
  moments of the icosphere; inside it tends to 1 like the geometry error, h^2.
- Laplace double layer of the density 1 (Gauss): -1 inside, 0 outside.
- Green's representation u = SL(du/dn) - DL(u) inside the sphere for the
  harmonic (Laplace) and radiating (Helmholtz) fields of an outside point
  source, with P0 projections of the traces: second-order convergence.
"""
import numpy as np
import pytest

import ngbem_amd as nb
from oracle import oracle as O


def _p0_traces(mesh, u_and_grad):
    rule = O.trig_rule(6)
    bary = np.stack([rule[:, 0], rule[:, 1], 1 - rule[:, 0] - rule[:, 1]], 1)
    X = np.einsum("qk,tkc->tqc", bary, mesh.xyz[mesh.tri])
    w = rule[:, 2] * 2
    n = mesh.normals()
    u, g = u_and_grad(X)
    dn = (g * n[:, None, :]).sum(-1)
    return (w[None] * u).sum(1) / w.sum(), (w[None] * dn).sum(1) / w.sum()


def test_single_and_double_layer_of_one():
    m = nb.icosphere(3)
    pts = np.array([[0, 0, 3.0], [2.0, 1.0, -1.0], [0.1, 0.2, -0.1], [0.0, 0.0, 0.0]])
    sl = O.OracleOp(m, O.LAPLACE_SL, 0.0, O.P0, O.P0, intorder=4)
    dl = O.OracleOp(m, O.LAPLACE_DL, 0.0, O.P0, O.P0, intorder=4)
    u = sl.potential(np.ones(m.nt), pts)[:, 0]
    r = np.linalg.norm(pts[:2], axis=1)
    assert np.allclose(u[:2].real * 4 * np.pi * r / m.areas().sum(), 1.0, atol=1e-6)
    assert np.allclose(u[2:].real, 1.0, atol=3e-3) and np.all(u[2:].real < 1.0)
    assert np.abs(u.imag).max() == 0.0
    d = dl.potential(np.ones(m.nt), pts)[:, 0].real
    assert np.abs(d[:2]).max() < 1e-10 and np.abs(d[2:] + 1.0).max() < 1e-7  # quadrature error only


@pytest.mark.parametrize("kappa", [0.0, 2.0])
def test_green_representation_converges(kappa):
    x0 = np.array([0.3, -0.2, 2.5])
    ip = np.array([[0.1, 0.2, -0.1], [0.3, -0.4, 0.2], [-0.5, 0.1, 0.1]])

    def field(X):
        D = X - x0
        R = np.linalg.norm(D, axis=-1)
        u = np.exp(1j * kappa * R) / (4 * np.pi * R)
        du = u * (1j * kappa - 1.0 / R)  # d/dR
        return u, du[..., None] * D / R[..., None]

    exact = field(ip)[0]
    errs = []
    for lvl in (2, 3):
        m = nb.icosphere(lvl)
        uP, dnP = _p0_traces(m, field)
        if kappa == 0.0:
            sl = O.OracleOp(m, O.LAPLACE_SL, 0.0, O.P0, O.P0, intorder=4)
            dl = O.OracleOp(m, O.LAPLACE_DL, 0.0, O.P0, O.P0, intorder=4)
        else:
            sl = O.OracleOp(m, O.HELMHOLTZ_SL, kappa, O.P0, O.P0, intorder=4)
            dl = O.OracleOp(m, O.HELMHOLTZ_DL, kappa, O.P0, O.P0, intorder=4)
        rep = sl.potential(dnP, ip)[:, 0] - dl.potential(uP, ip)[:, 0]
        errs.append(np.abs(rep - exact).max() / np.abs(exact).max())
    # measured: kappa 0: 4.0e-4, 9.9e-5 (L2, L3); kappa 2: 4.8e-3, 1.2e-3
    assert errs[1] < (2e-4 if kappa == 0.0 else 2e-3) and errs[1] < 0.3 * errs[0]


def test_potential_dimension_follows_the_trial_evaluator():
    m = nb.icosphere(1)
    hs = O.OracleOp(m, O.HELMHOLTZ_HS, 2.0, O.P1, O.P1, intorder=3)
    assert hs.potential(np.ones(m.nv), [[0, 0, 2.0]]).shape == (1, 4)
    msl = O.OracleOp(m, O.MAXWELL_SL, 2.0, O.HDIV, O.HDIV, intorder=3)
    assert msl.potential(np.ones(msl.ntrial), [[0, 0, 2.0]]).shape == (1, 4)
