"""Convergence of the oracle's higher-order operators to an exact solution on
the box (flat triangles exact): the Neumann trace of 1/|x - x0| from
V t = (1/2 M + K) g with H1 order p / SurfaceL2 order p - 1 converges at
rate ~p in h, and each order is several times more accurate than the one
below.  Measured (n = 2, 4): p = 1: 0.187, 0.093; p = 2: 0.041, 0.012;
p = 3: 0.0073, 0.0015."""
import numpy as np
import pytest

from oracle import oracle as O
import box_laplace as B


def _error(n, p):
    m, h1, l2 = B.spaces(n, p)
    io = 2 * p + 4
    V = O.OracleOp(m, O.LAPLACE_SL, 0.0, trial=l2.ikind, test=l2.ikind, intorder=io).dense().real
    K = O.OracleOp(m, O.LAPLACE_DL, 0.0, trial=h1.ikind, test=l2.ikind, intorder=io).dense().real
    return B.neumann_error(m, h1, l2, V, K)[0]


def test_higher_orders_converge_faster():
    e = {(n, p): _error(n, p) for n in (2, 4) for p in (1, 2, 3)}
    rate = {p: np.log2(e[(2, p)] / e[(4, p)]) for p in (1, 2, 3)}
    assert rate[1] > 0.9 and rate[2] > 1.6 and rate[3] > 2.0, rate
    for n in (2, 4):
        assert e[(n, 2)] < e[(n, 1)] / 3 and e[(n, 3)] < e[(n, 2)] / 3, e
