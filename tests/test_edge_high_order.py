"""Edge spaces of orders 2-4 and H1 order 4: the basis, CPU only.

The product's basis (device evaluators, written as weights on the curls c_k =
n x grad_Gamma lambda_k with div = (dg/dlambda_(c+1) - dg/dlambda_(c+2)) / J,
mirrored in numpy by ngbem_amd.spaces) against the oracle's restatement
through the reference element (NGSolve's Piola maps, DiffOpMaxwellNew /
DiffOpMaxwell), and what the mathematics fixes:
- the Piola path reproduces the oracle's pinned order 0-1 functions;
- product and oracle agree at every order, flat and curved;
- order p spans all tangential fields of degree p (rank (p+1)(p+2) of the
  fitted coefficients, zero residual against complete polynomials);
- HDivSurface: normal traces continuous across every edge, zero for the
  interior functions; HCurl: tangential traces continuous;
- the divergence theorem on every element (flat and curved), which pins the
  scalar parts (div phi, curl N);
- H1 order 4: the 15 functions span P4 and are continuous across edges.
"""
import numpy as np
import pytest

import ngbem_amd as nb
from oracle import oracle as O

HDIV = (O.HDIV, O.HDIV1, O.HDIV2, O.HDIV3, O.HDIV4)
HCURL = (O.HCURL, O.HCURL1, O.HCURL2, O.HCURL3, O.HCURL4)


def _pts(n, seed=0):
    u = np.random.default_rng(seed).random((n, 2))
    xi, eta = u[:, 0] * (1 - u[:, 1]), u[:, 1]
    return np.stack([xi, eta, 1 - xi - eta], 1)


@pytest.mark.parametrize("cls,kinds", [(nb.HDivSurface, HDIV), (nb.HCurl, HCURL)], ids=["HDiv", "HCurl"])
@pytest.mark.parametrize("curved", [False, True])
def test_product_basis_equals_oracle(cls, kinds, curved):
    m = nb.icosphere(1)
    if curved:
        m.Curve(3)
    bary = _pts(5)
    for p in range(5):
        sp, osp = cls(m, order=p), O.OracleSpace(m, kinds[p])
        assert np.array_equal(osp.elem_dofs, sp.elem_dofs)
        tr, dv = sp.trace_shapes(bary), sp.div_shapes(bary)
        for e in (0, 11, 57):
            for q in range(len(bary)):
                o = osp.eval(O.EV_MAXW, e, bary[q, 0], bary[q, 1])
                ref = np.concatenate([tr[e, q], dv[e, q][:, None]], 1)
                assert np.abs(o - ref).max() <= (1e-13 if curved else 1e-15) * np.abs(ref).max(), (p, e, q)
                if p <= 1 and not curved:  # the reference-element path against the pinned one
                    op = osp.eval(O.EV_MAXW, e, bary[q, 0], bary[q, 1], piola=True)
                    assert np.abs(op - o).max() <= 1e-15 * np.abs(o).max()


def _tangent_coeffs(m, e, vec):
    """Coefficients (alpha, beta) of tangent fields vec (q x nloc x 3) of the
    flat element e in the basis t1 = p0 - p2, t2 = p1 - p2."""
    p0, p1, p2 = m.xyz[m.tri[e]]
    T = np.stack([p0 - p2, p1 - p2], 1)  # 3 x 2
    return np.einsum("ij,qlj->qli", np.linalg.pinv(T), vec)


@pytest.mark.parametrize("cls", [nb.HDivSurface, nb.HCurl], ids=["HDiv", "HCurl"])
def test_spans_all_fields_of_degree_p(cls):
    m = nb.icosphere(1)
    bary = _pts(60, seed=1)
    for p in range(5):
        sp = cls(m, order=p)
        C = _tangent_coeffs(m, 3, sp.vec_shapes(bary)[3])  # q x nloc x 2
        deg = max(p, 1)  # RT0 / Whitney: degree 1
        V = np.stack([bary[:, 0] ** a * bary[:, 1] ** b for d in range(deg + 1) for b in range(d + 1)
                      for a in [d - b]], 1)
        for comp in range(2):
            coef, res, *_ = np.linalg.lstsq(V, C[:, :, comp], rcond=None)
            fit = V @ coef
            assert np.abs(fit - C[:, :, comp]).max() <= 1e-12 * np.abs(C).max(), p
        nloc = sp.elem_dofs.shape[1]
        assert np.linalg.matrix_rank(C.transpose(1, 0, 2).reshape(nloc, -1), tol=1e-10) == nloc
        if p >= 1:
            assert nloc == (p + 1) * (p + 2)  # = dim of all degree-p tangential fields


def _edge_points(l, tau):
    """Barycentrics of local edge l = (l, l+1) at parameter tau (0 at v_l)."""
    b = np.zeros((len(tau), 3))
    b[:, l] = 1 - tau
    b[:, (l + 1) % 3] = tau
    return b


@pytest.mark.parametrize("curved", [False, True])
@pytest.mark.parametrize("cls", [nb.HDivSurface, nb.HCurl], ids=["HDiv", "HCurl"])
def test_traces_continuous_across_edges(cls, curved):
    m = nb.icosphere(1)
    if curved:
        m.Curve(2)
    tau = np.array([0.13, 0.5, 0.87])  # symmetric: reversed, the same points
    for p in (2, 3, 4):
        sp = cls(m, order=p)
        npe = p + 1
        vals, conorm = {}, {}
        for l in range(3):
            b = _edge_points(l, tau)
            X, n, J, t1, t2 = m.frames(b[:, 0], b[:, 1])
            # dX/dtau along the edge in the element's direction: X depends on
            # (lambda_0, lambda_1) = (xi, eta) with dX/dxi = t1, dX/deta = t2
            d = [t1, t2, 0.0 * t1]
            dX = d[(l + 1) % 3] - d[l]
            phi = sp.vec_shapes(b)
            for e in range(m.nt):
                a, c = m.tri[e, l], m.tri[e, (l + 1) % 3]
                key = (min(a, c), max(a, c))
                sgn = 1.0 if a < c else -1.0
                tg = sgn * dX[e]                       # tangent from the low to the high vertex
                order = slice(None) if a < c else slice(None, None, -1)
                if cls is nb.HDivSurface:              # normal flux with the co-normal t x n
                    tr = np.einsum("qlc,qc->ql", phi[e], np.cross(tg, n[e]))[order]
                else:                                  # tangential trace
                    tr = np.einsum("qlc,qc->ql", phi[e], tg)[order]
                dofs = sp.elem_dofs[e]
                own = range(l * npe, (l + 1) * npe)
                vals.setdefault(key, []).append({int(dofs[k]): tr[:, k] for k in own})
                # the other edges' and the interior functions: zero normal
                # (HDiv) / tangential (HCurl) trace on this edge
                others = [k for k in range(sp.elem_dofs.shape[1]) if k not in own]
                assert np.abs(tr[:, others]).max() <= 1e-12 * np.abs(tr).max()
        for key, sides in vals.items():
            assert len(sides) == 2
            for d0, v0 in sides[0].items():
                assert d0 in sides[1]
                scale = max(np.abs(v0).max(), 1e-300)
                assert np.abs(v0 - sides[1][d0]).max() <= 1e-10 * scale, (p, key, d0)


@pytest.mark.parametrize("curved", [False, True])
def test_divergence_theorem(curved):
    """int_T div phi dA = boundary integral of phi . nu ds per element, with
    nu ds = dX/dtau x n dtau on the counterclockwise boundary; HCurl: Stokes,
    int curl N dA = boundary integral of N . dX."""
    m = nb.icosphere(1)
    if curved:
        m.Curve(3)
    g, gw = np.polynomial.legendre.leggauss(12)
    g, gw = 0.5 * (g + 1), 0.5 * gw
    U, Vv = np.meshgrid(g, g, indexing="ij")
    W = ((gw[:, None] * gw[None, :]) * (1 - U)).ravel()
    xs, ys = U.ravel(), (Vv * (1 - U)).ravel()
    bary = np.stack([xs, ys, 1 - xs - ys], 1)
    J = m.geom(xs, ys)[2]
    for p in (0, 2, 3, 4):
        for cls in (nb.HDivSurface, nb.HCurl):
            sp = cls(m, order=p)
            area_int = np.einsum("q,tq,tql->tl", W, J, sp.div_shapes(bary))
            bnd = np.zeros_like(area_int)
            for l in range(3):  # counterclockwise: v0 -> v1 -> v2 -> v0 is edges 0, 1, 2
                b = _edge_points(l, g)
                X, n, Jb, t1, t2 = m.frames(b[:, 0], b[:, 1])
                d = [t1, t2, 0.0 * t1]
                dX = d[(l + 1) % 3] - d[l]
                phi = sp.vec_shapes(b)
                if cls is nb.HDivSurface:
                    f = np.einsum("tqlc,tqc->tql", phi, np.cross(dX, n))
                else:
                    f = np.einsum("tqlc,tqc->tql", phi, dX)
                bnd += np.einsum("q,tql->tl", gw, f)
            err = np.abs(area_int - bnd).max() / np.abs(bnd).max()
            assert err <= (1e-12 if not curved else 1e-9), (cls.__name__, p, err)


def test_h1_order4_span_and_continuity():
    m = nb.icosphere(1)
    h = nb.H1(m, order=4)
    oh = O.OracleSpace(m, O.P4)
    assert np.array_equal(oh.elem_dofs, h.elem_dofs) and h.elem_dofs.shape[1] == 15
    bary = _pts(40, seed=2)
    S = h.shapes(bary) * 1.0
    # spans P4 on the triangle
    V = np.stack([bary[:, 0] ** a * bary[:, 1] ** b for d in range(5) for b in range(d + 1) for a in [d - b]], 1)
    coef = np.linalg.lstsq(V, S, rcond=None)[0]
    assert np.abs(V @ coef - S).max() <= 1e-12
    assert np.linalg.matrix_rank(coef, tol=1e-10) == 15
    # oracle values and rotated gradients equal the product's functions
    for e in (0, 9):
        for q in range(5):
            o = oh.eval(O.EV_ID, e, bary[q, 0], bary[q, 1])[:, 0]
            assert np.abs(o - S[q] * h.elem_signs()[e]).max() <= 1e-15
    # continuity across every edge (values at three points per edge)
    tau = np.array([0.1, 0.5, 0.9])  # symmetric: reversed, the same points
    seen = {}
    for l in range(3):
        Sl = h.shapes(_edge_points(l, tau)) * 1.0
        for e in range(m.nt):
            a, c = m.tri[e, l], m.tri[e, (l + 1) % 3]
            vals = (Sl * h.elem_signs()[e])[:: 1 if a < c else -1]
            u = np.zeros((len(tau), h.ndof))
            u[:, h.elem_dofs[e]] = vals
            key = (min(a, c), max(a, c))
            if key in seen:
                assert np.abs(seen[key] - u).max() <= 1e-14
            else:
                seen[key] = u
