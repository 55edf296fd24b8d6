"""Pins for the oracle's quadrature rules (src/intrules.cpp:13-136 and the
triangle / Gauss rules that replace NGSolve's).

- Gauss-Legendre: exact moments and agreement with numpy's leggauss.
- Sauter-Schwab rules: weights sum to |T|^2 = 1/4; polynomials integrate exactly.
- Triangle rules: exact for all monomials up to their degree.
- Singular integrals of 1/|x-y| over identical, edge- and vertex-sharing
  triangle pairs against the closed-form potential of a triangle (Wilton et
  al. 1984) integrated over the outer triangle with an independent rule.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_gauss_legendre_matches_numpy():
    for n in range(1, 16):
        x, w = O.gauss01(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.allclose(x, 0.5 * (xr + 1), atol=1e-15, rtol=0)
        assert np.allclose(w, 0.5 * wr, atol=1e-15, rtol=0)
        for k in range(2 * n):
            assert abs(np.dot(w, x ** k) - 1.0 / (k + 1)) < 1e-14


def tri_moment(a, b):
    """int_T x^a y^b over the unit right triangle."""
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("order", [2, 3, 4, 6, 10])
def test_sauter_schwab_weight_sum(kind, order):
    r = O.rule4(kind, order)
    n = order // 2 + 1
    assert len(r) == [6, 5, 2][kind] * n ** 4
    assert abs(r[:, 4].sum() - 0.25) < 1e-14
    # all points inside the unit right triangle
    for c in (0, 2):
        assert (r[:, c] >= -1e-15).all() and (r[:, c + 1] >= -1e-15).all()
        assert (r[:, c] + r[:, c + 1] <= 1 + 1e-15).all()


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_sauter_schwab_polynomial_exactness(kind):
    order = 10  # 6-point Gauss in each Duffy variable: exact to total degree 8
    r = O.rule4(kind, order)
    x0, x1, y0, y1, w = r.T
    rng = np.random.default_rng(kind)
    for _ in range(40):
        a, b, c, d = rng.integers(0, 3, size=4)
        if a + b + c + d > 8:
            continue
        exact = tri_moment(a, b) * tri_moment(c, d)
        got = np.sum(w * x0 ** a * x1 ** b * y0 ** c * y1 ** d)
        assert abs(got - exact) <= 1e-13 * max(exact, 1e-3), (a, b, c, d)


@pytest.mark.parametrize("order,degree", [(1, 1), (2, 2), (3, 4), (4, 4), (5, 5), (6, 6), (7, 7), (9, 9), (12, 12)])
def test_triangle_rule_exactness(order, degree):
    t = O.trig_rule(order)
    xi, eta, w = t.T
    assert abs(w.sum() - 0.5) < 1e-15
    assert (w > 0).all()
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            assert abs(np.sum(w * xi ** a * eta ** b) - tri_moment(a, b)) < 1e-15, (a, b)
    if order in (3, 4):
        assert len(t) == 6  # Dunavant degree 4
    if order == 5:
        assert len(t) == 7


# ---- singular integrals against the analytic potential of a triangle ----

def triangle_potential(P, V):
    """int_T 1/|P - y| dA(y) for the flat triangle V (3x3), closed form of
    Wilton, Rao, Glisson, Schaubert, Al-Bundak, Butler (IEEE TAP 1984)."""
    n = np.cross(V[1] - V[0], V[2] - V[0])
    n /= np.linalg.norm(n)
    d = np.dot(P - V[0], n)
    rho = P - d * n
    total = 0.0
    for i in range(3):
        a, b = V[i], V[(i + 1) % 3]
        lt = (b - a) / np.linalg.norm(b - a)
        u = np.cross(lt, n)  # outward in-plane edge normal (CCW triangle)
        lm = np.dot(a - rho, lt)
        lp = np.dot(b - rho, lt)
        p0 = np.dot(a - rho, u)
        R0sq = p0 * p0 + d * d
        Rm = math.sqrt(R0sq + lm * lm)
        Rp = math.sqrt(R0sq + lp * lp)
        if abs(p0) < 1e-15:
            continue
        # p0 * ln((Rp + lp)/(Rm + lm)), written stably for lm, lp < 0
        if lm >= 0 or lp >= 0:
            log_term = math.log((Rp + lp) / (Rm + lm))
        else:
            log_term = math.log((Rm - lm) / (Rp - lp))
        ad = abs(d)
        atan_term = math.atan2(p0 * lp, R0sq + ad * Rp) - math.atan2(p0 * lm, R0sq + ad * Rm)
        total += p0 * log_term - ad * atan_term
    return total


def outer_rule(n):
    """collapsed Gauss rule on the unit right triangle (numpy's leggauss)."""
    g, gw = np.polynomial.legendre.leggauss(n)
    g = 0.5 * (g + 1)
    gw = 0.5 * gw
    pts, wts = [], []
    for i in range(n):
        for j in range(n):
            pts.append((g[i], (1 - g[i]) * g[j]))
            wts.append(gw[i] * gw[j] * (1 - g[i]))
    return np.array(pts), np.array(wts)


def pair_integral_analytic(T1, T2, n=48):
    pts, wts = outer_rule(n)
    J1 = np.linalg.norm(np.cross(T1[0] - T1[2], T1[1] - T1[2]))
    s = 0.0
    for (xi, eta), w in zip(pts, wts):
        X = xi * T1[0] + eta * T1[1] + (1 - xi - eta) * T1[2]
        s += w * triangle_potential(X, T2)
    return s * J1


def pair_integral_ss(T1, T2, kind, order, px, py):
    """Sauter-Schwab integral of 1/|x-y| with the same vertex permutations the
    oracle uses: SS vertex k of T1 is T1[px[k]]."""
    r = O.rule4(kind, order)
    x0, x1, y0, y1, w = r.T
    A = T1[list(px)]
    B = T2[list(py)]
    lx = np.stack([1 - x0 - x1, x0, x1], 1)
    ly = np.stack([1 - y0 - y1, y0, y1], 1)
    X = lx @ A
    Y = ly @ B
    J1 = np.linalg.norm(np.cross(T1[0] - T1[2], T1[1] - T1[2]))
    J2 = np.linalg.norm(np.cross(T2[0] - T2[2], T2[1] - T2[2]))
    return J1 * J2 * np.sum(w / np.linalg.norm(X - Y, axis=1))


def test_wilton_formula_regular_pair():
    # independent check of the analytic potential: far-apart triangles vs a
    # plain tensor Gauss rule
    T = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0.]])
    P = np.array([0.3, 0.2, 2.5])
    pts, wts = outer_rule(30)
    Y = pts[:, :1] * T[0] + pts[:, 1:] * T[1] + (1 - pts.sum(1))[:, None] * T[2]
    brute = np.sum(wts / np.linalg.norm(P - Y, axis=1)) * 1.0
    assert abs(brute - triangle_potential(P, T)) < 1e-12


TRI = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 0.9, 0.1]])


@pytest.mark.parametrize("case", ["identical", "edge_coplanar", "edge_bent", "vertex_bent"])
def test_singular_rules_vs_analytic(case):
    T1 = TRI.copy()
    if case == "identical":
        T2 = T1.copy()
        kind, px, py = 0, (2, 0, 1), (2, 0, 1)
    elif case.startswith("edge"):
        # share edge T1[0]-T1[1]; T2 runs through it in the opposite direction
        apex = np.array([0.6, -0.8, 0.0 if case == "edge_coplanar" else 0.5])
        T2 = np.array([T1[1], T1[0], apex])
        kind = 1
        px, py = (0, 1, 2), (1, 0, 2)  # SS origin = T1[0] = T2[1]
    else:
        T2 = np.array([T1[0], [-0.7, -0.2, 0.4], [-0.3, -0.9, -0.2]])
        kind, px, py = 2, (0, 1, 2), (0, 1, 2)
    exact = pair_integral_analytic(T1, T2)
    errs = [abs(pair_integral_ss(T1, T2, kind, o, px, py) - exact) / abs(exact) for o in (4, 8, 14)]
    assert errs[-1] < 2e-7, errs
    assert errs[1] < errs[0] or errs[0] < 1e-9
