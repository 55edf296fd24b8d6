"""Quadrature convergence in the integration order, as the reference's
tests/test_DLIntegration.py checks it (CPU oracle).

The reference assembles the Laplace double layer H1 p1 -> SurfaceL2 p0
densely at intorder 2, 4, ..., 20 and requires the spectral norm of the
difference to the intorder-26 matrix (30 power steps of D^T D) to stay
below 5x its table (tests/test_DLIntegration.py:7-16, 37-45; a Netgen sphere
mesh with maxh 0.5, flat elements).  Here: the same computation on the
icosphere of level 1 (80 flat triangles).  Besides the reference's bound,
every entry lies within a factor 0.15-2 of the reference's own value: the
decay per order is that of the reference's rules (Sauter-Schwab with the
Gauss-Legendre rule of n = intorder/2 + 1 points, and the triangle rule of
the same order), which a rule with fewer points or a wrong weight would not
reproduce.
"""
import numpy as np

import ngbem_amd as nb
from oracle import oracle as O

# tests/test_DLIntegration.py:7-16
DL_ERRREF = [0.00014356401693177312, 1.1556523998207007e-05, 9.959433217687845e-07, 9.474259269075394e-08,
             9.6163481617507e-09, 9.182994523208552e-10, 8.329321209219403e-11, 7.326096658552548e-12,
             8.665555899250029e-13, 1.6241027931801214e-13]


def spectral_error_dtd(D, seed=1, steps=30):
    """||D x|| after `steps` power steps of D^T D (tests/test_DLIntegration.py:19-25)."""
    z = np.random.default_rng(seed).random(D.shape[1])
    for _ in range(steps):
        x = z / np.linalg.norm(z)
        y = D @ x
        z = D.T @ y
    return np.linalg.norm(y)


def test_dl_integration_order_table():
    m = nb.icosphere(1)
    dl = lambda io: O.OracleOp(m, O.LAPLACE_DL, 0.0, trial=O.P1, test=O.P0, intorder=io).dense().real  # noqa: E731
    ref = dl(26)
    errs = np.array([spectral_error_dtd(ref - dl(2 * (n + 1))) for n in range(10)])
    print("intorder 2..20:", errs, errs / DL_ERRREF)
    assert np.all(errs < 5 * np.array(DL_ERRREF))
    ratio = errs[:9] / np.array(DL_ERRREF[:9])
    assert np.all((ratio > 0.15) & (ratio < 2.0))
