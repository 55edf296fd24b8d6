"""Edge spaces of orders 2-4, edge spaces on curved meshes and H1 order 4 on
the GPU (the wide evaluators and k_near_wide), through the C-ABI, against
the oracle (reference-element restatement of the bases), and the
reference's tests/test_Maxwell_Mie.py at its orders.

Bars: near-field entries within 1e-12 of the oracle's (max over the block's
largest entry), at the GPU's block dofs; low-rank blocks against the exact
block of the same dofs within 100 eps or 10x the oracle's own cross
approximation error; the H-matrix apply against the oracle's dense operator
within 10 eps relative to ||A|| ||x||; the potential within 1e-12.
"""
import numpy as np
import pytest
import torch

import ngbem_amd as nb
from oracle import oracle as O

pytestmark = pytest.mark.gpu

EDGE = {("d", 0): O.HDIV, ("d", 1): O.HDIV1, ("d", 2): O.HDIV2, ("d", 3): O.HDIV3, ("d", 4): O.HDIV4,
        ("c", 0): O.HCURL, ("c", 1): O.HCURL1, ("c", 2): O.HCURL2, ("c", 3): O.HCURL3, ("c", 4): O.HCURL4,
        ("h", 1): O.P1, ("h", 2): O.P2, ("h", 3): O.P3, ("h", 4): O.P4, ("l", 0): O.P0, ("l", 3): O.L2_3}
CLS = {"d": nb.HDivSurface, "c": nb.HCurl, "h": nb.H1, "l": nb.SurfaceL2}

# (kernel, trial (family, order), test (family, order), level, geometry order, kappa, leafsize)
CASES = [
    ("MSL", ("d", 2), ("d", 2), 1, 1, 3.0, 24),
    ("MSL", ("d", 4), ("d", 4), 1, 1, 3.0, 60),
    ("MSLC", ("c", 3), ("c", 3), 1, 1, 3.0, 40),
    ("MDL", ("c", 3), ("d", 3), 1, 1, 3.0, 40),
    ("MDL", ("c", 2), ("d", 4), 1, 1, 3.0, 40),
    ("MSL", ("d", 0), ("d", 0), 2, 2, 3.0, 16),
    ("MSL", ("d", 2), ("d", 2), 1, 3, 3.0, 24),
    ("MDL", ("c", 1), ("d", 1), 2, 2, 3.0, 24),
    ("LHS", ("h", 4), ("h", 4), 2, 1, 0.0, 60),
    ("HHS", ("h", 4), ("h", 4), 1, 1, 3.0, 40),
    ("LDL", ("h", 4), ("l", 3), 2, 1, 0.0, 60),
    ("HDL", ("h", 4), ("h", 4), 1, 4, 3.0, 40),
]


def _build(case, eps=1e-8):
    name, (ft, ot), (fs, os_), level, k, kappa, leaf = case
    mesh = nb.icosphere(level)
    if k > 1:
        mesh.Curve(k)
    tr, te = CLS[ft](mesh, order=ot), CLS[fs](mesh, order=os_)
    kw = dict(intorder=4, leafsize=leaf, eta=2.0, eps=eps)
    print(f"wide case {case}: {tr.ndof} x {te.ndof} dofs", flush=True)
    if name == "MSL":
        op, kid = nb.MaxwellSingleLayerPotentialOperator(tr, kappa, **kw), O.MAXWELL_SL
    elif name == "MSLC":
        op, kid = nb.MaxwellSingleLayerPotentialOperatorCurl(tr, kappa, **kw), O.MAXWELL_SL
    elif name == "MDL":
        op, kid = nb.MaxwellDoubleLayerPotentialOperator(tr, te, kappa, **kw), O.MAXWELL_DL
    elif name == "LHS":
        op, kid = nb.HypersingularOperator(tr, **kw), O.LAPLACE_HS
    elif name == "HHS":
        op, kid = nb.HelmholtzHypersingularOperator(tr, te, kappa=kappa, **kw), O.HELMHOLTZ_HS
    elif name == "LDL":
        op, kid = nb.DoubleLayerPotentialOperator(tr, te, **kw), O.LAPLACE_DL
    else:
        op, kid = nb.HelmholtzDoubleLayerPotentialOperator(tr, te, kappa=kappa, **kw), O.HELMHOLTZ_DL
    ref = O.OracleOp(mesh, kid, kappa=kappa, trial=EDGE[(ft, ot)], test=EDGE[(fs, os_)], **kw)
    assert (ref.ntrial, ref.ntest) == (tr.ndof, te.ndof)
    return op, ref


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[1][0]}{c[1][1]}{c[2][0]}{c[2][1]}-L{c[3]}-k{c[4]}"
                                             for c in CASES])
def test_wide_blocks_and_apply(gpu, case):
    eps = 1e-8
    op, ref = _build(case, eps)
    nfar, worst = 0, 0.0
    for b in range(op.num_blocks()):
        rows, cols = op.block_dofs(b)
        data = op.block_data(b)
        D = ref.dense(cols, rows)
        if not op.is_complex:
            D = D.real
        if isinstance(data, tuple):
            nfar += 1
            UV = ref.aca(cols, rows)
            err, err_ref = np.linalg.norm(data[0] @ data[1] - D), np.linalg.norm(UV[0] @ UV[1] - D)
            bar = max(100 * eps * np.linalg.norm(D), 10 * err_ref)
            worst = max(worst, err / bar)
            assert err <= bar, f"far block {b}: {err} vs {err_ref}"
        else:
            assert np.abs(data - D).max() <= 1e-12 * max(np.abs(D).max(), 1e-300), f"block {b}"
    print(f"{op.num_blocks()} blocks, {nfar} low rank (largest error / bar {worst:.2f})", flush=True)
    A = ref.dense()
    if not op.is_complex:
        A = A.real
    rng = np.random.default_rng(3)
    for trans in (False, True):
        n_in = A.shape[0] if trans else A.shape[1]
        x = rng.standard_normal(n_in) + (1j * rng.standard_normal(n_in) if op.is_complex else 0)
        y = ((op.mat.T if trans else op.mat) * torch.from_numpy(x).to(gpu)).cpu().numpy()
        y_ref = (A.T if trans else A) @ x
        assert np.linalg.norm(y - y_ref) <= 10 * eps * np.linalg.norm(A) * np.linalg.norm(x)


@pytest.mark.parametrize("which", ["MSL-d3", "MDL-c2-curved", "HDL-h4"])
def test_wide_potential(gpu, which):
    case = {"MSL-d3": ("MSL", ("d", 3), ("d", 3), 1, 1, 3.0, 40),
            "MDL-c2-curved": ("MDL", ("c", 2), ("d", 2), 1, 3, 3.0, 40),
            "HDL-h4": ("HDL", ("h", 4), ("h", 4), 1, 1, 3.0, 40)}[which]
    op, ref = _build(case)
    rng = np.random.default_rng(9)
    c = rng.standard_normal(op.trial_space.ndof) + 1j * rng.standard_normal(op.trial_space.ndof)
    pts = rng.normal(size=(200, 3)) * 2.0
    pts = pts[np.linalg.norm(pts, axis=1) > 1.3]
    nrm = pts / np.linalg.norm(pts, axis=1)[:, None]
    pot = op.GetPotential(torch.from_numpy(c).to(gpu))
    u = pot(torch.from_numpy(pts).to(gpu), torch.from_numpy(nrm).to(gpu)).cpu().numpy().reshape(len(pts), -1)
    u_ref = ref.potential(c, pts, nrm).reshape(len(pts), -1)
    assert u.shape == u_ref.shape
    assert np.abs(u - u_ref).max() <= 1e-12 * np.abs(u_ref).max()


def _surface_rule(mesh, order):
    """Points, unit normals and weights J w of a triangle rule on every element
    (flat or curved)."""
    r = O.trig_rule(order)
    bary = np.stack([r[:, 0], r[:, 1], 1 - r[:, 0] - r[:, 1]], 1)
    X, n, J = mesh.geom(r[:, 0], r[:, 1])
    return bary, X, n, r[:, 2][None, :] * J


def _mie_efie(level, order, kappa, radius, intorder, leafsize, eta, eps, gpu, tol=1e-8):
    """The reference's tests/test_Maxwell_Mie.py on an icosphere of the given
    level and radius, Curve(4): HDivSurface of the order, rhs -int E_inc . v
    with E_inc = (1, 0, 0) exp(i kappa z) (bonus integration order), GMRES
    preconditioned by the mass inverse; returns the reference's error
    sqrt(int |kappa j - n x curl E|^2) and the dof count."""
    import mie_maxwell as M
    import scipy.sparse as sps
    from scipy.sparse.linalg import splu

    mesh = nb.icosphere(level, radius=radius).Curve(4)
    hd = nb.HDivSurface(mesh, order=order)
    bary, X, n, w = _surface_rule(mesh, 2 * order + 10)
    phi = hd.vec_shapes(bary)
    E = np.zeros(X.shape, complex)
    E[..., 0] = np.exp(1j * kappa * X[..., 2])
    b = np.zeros(hd.ndof, complex)
    np.add.at(b, hd.elem_dofs, -np.einsum("tq,tqc,tqlc->tl", w, E, phi))
    r, c, v = nb.spaces.mass_matrix_coo(hd, hd)
    lu = splu(sps.csc_matrix((v.astype(complex), (r, c)), shape=(hd.ndof, hd.ndof)))
    pre = lambda x: torch.from_numpy(lu.solve(x.cpu().numpy().astype(complex))).to(gpu)  # noqa: E731
    V = nb.MaxwellSingleLayerPotentialOperator(hd, kappa, intorder=intorder, leafsize=leafsize, eta=eta, eps=eps)
    j = nb.GMRes(V.mat, torch.from_numpy(b).to(gpu), pre=pre, tol=tol, maxsteps=1000, printrates=False).cpu().numpy()
    jh = np.einsum("tl,tqlc->tqc", j[hd.elem_dofs], phi)
    Xs = radius * X / np.linalg.norm(X, axis=-1)[..., None]
    jex = M.current(Xs.reshape(-1, 3), kappa, radius).reshape(X.shape)
    return kappa * np.sqrt(np.sum(w * (np.abs(jh - jex) ** 2).sum(-1))), hd.ndof


def test_reference_maxwell_mie_order4(gpu):
    """tests/test_Maxwell_Mie.py: sphere of radius 0.25 curved to order 4,
    HDivSurface order 4, kappa 5, intorder 16, leafsize 120, eta 3, eps 1e-8,
    GMRES to 1e-8: L2 error of the current below the reference's 0.0015.  Its
    mesh is Netgen's at maxh 1, whose size the sphere's curvature sets; here
    icosphere levels 0-2 (20, 80, 320 triangles): the error falls at the
    O(h^(p+1)) rate of degree-4 fields (32x per level; at least 16x asserted)
    and is below the reference's bar at level 2."""
    errs = {}
    for level in (0, 1, 2):
        errs[level], ndof = _mie_efie(level, 4, 5.0, 0.25, 16, 120, 3.0, 1e-8, gpu)
        print(f"Maxwell Mie, HDivSurface 4, Curve(4), icosphere {level}: {ndof} dofs, error {errs[level]:.3e}",
              flush=True)
    assert errs[2] < 0.0015
    assert errs[0] / errs[1] > 16 and errs[1] / errs[2] > 16


def test_reference_maxwell_mie_convergence_in_order(gpu):
    """tests/Maxwell_Mie_Convergence.py at one mesh (icosphere level 1, radius
    0.25, Curve(4)), orders 0-4 with intorder 8 + 2 order, leafsize 320, eta 3,
    eps 1e-10, GMRES to 1e-11: the error falls with the order (p-refinement,
    each order at least 3x below the previous)."""
    errs = []
    for order in range(5):
        err, ndof = _mie_efie(1, order, 5.0, 0.25, 8 + 2 * order, 320, 3.0, 1e-10, gpu, tol=1e-11)
        errs.append(err)
        print(f"order {order}: {ndof} dofs, error {err:.3e}", flush=True)
    errs = np.array(errs)
    assert np.all(errs[1:] < errs[:-1] / 3.0), errs


def test_reference_maxwell_box_order3(gpu):
    """tests/test_Maxwell_Box.py: the field of a dipole at the origin (kappa =
    1 / h = 1, direction (1.192, -0.189, 2.745)) outside the box [-5, 5] x
    [-1, 1]^2 (the reference fuses two smaller boxes onto it; here the main
    box alone, 10 x 2 x 2 cells of size 1), HCurl / HDivSurface order 3:
    Dirichlet data m = the H(curl) projection of -(n x E) x n onto HCurl
    (inner product (m, N) + (curl m, curl N), curl_Gamma m = -n . curl E; like
    NGSolve's Set(dual=True) it controls the surface curl, which the plain L2
    projection does not: 6.2e-2 with it), V j = (-1/2 M + K) m by GMRES with
    the mass-inverse preconditioner (intorder 16, leafsize 320, eta 0: all
    blocks dense, eps 1e-8); the relative L2 error of j against n x curl E /
    kappa stays below the reference's 0.015 (measured 7.0e-3; the L2 best
    approximation of the current is 5.1e-3)."""
    import scipy.sparse as sps
    from scipy.sparse.linalg import splu
    from test_oracle_maxwell import dipole_field

    mesh = nb.box((10, 2, 2), size=(10.0, 2.0, 2.0))
    hc, hd = nb.HCurl(mesh, order=3), nb.HDivSurface(mesh, order=3)
    kappa = 1.0
    e = np.array([1.192, -0.189, 2.745]) / np.linalg.norm([1.192, -0.189, 2.745])
    bary, X, n, w = _surface_rule(mesh, 12)
    E, curlE = dipole_field(X, kappa, e)
    mex = -(E - (E * n).sum(-1)[..., None] * n)

    def mass(a, b):
        r, c, v = nb.spaces.mass_matrix_coo(a, b)
        return sps.csc_matrix((v, (r, c)), shape=(b.ndof, a.ndof))

    Nc, Cc = hc.vec_shapes(bary), hc.div_shapes(bary)  # N and curl N
    bm = np.zeros(hc.ndof, complex)
    np.add.at(bm, hc.elem_dofs, np.einsum("tq,tqc,tqlc->tl", w, mex, Nc))
    np.add.at(bm, hc.elem_dofs, np.einsum("tq,tq,tql->tl", w, -(curlE * n).sum(-1), Cc))
    d = hc.elem_dofs
    nl = d.shape[1]
    Cm = sps.csc_matrix((np.einsum("tq,tqi,tqj->tij", w, Cc, Cc).reshape(-1),
                         (np.repeat(d, nl, 1).reshape(-1), np.tile(d, (1, nl)).reshape(-1))), shape=(hc.ndof, hc.ndof))
    m = splu((mass(hc, hc) + Cm).astype(complex)).solve(bm)
    kw = dict(intorder=16, leafsize=320, eta=0.0, eps=1e-8)
    V = nb.MaxwellSingleLayerPotentialOperator(hd, kappa, **kw)
    K = nb.MaxwellDoubleLayerPotentialOperator(hc, hd, kappa, **kw)
    mt = torch.from_numpy(m).to(gpu)
    rhs = (K.mat * mt).cpu().numpy() - 0.5 * (mass(hc, hd) @ m)
    lu = splu(mass(hd, hd).astype(complex))
    pre = lambda x: torch.from_numpy(lu.solve(x.cpu().numpy().astype(complex))).to(gpu)  # noqa: E731
    j = nb.GMRes(V.mat, torch.from_numpy(rhs).to(gpu), pre=pre, tol=1e-8, maxsteps=1000,
                 printrates=False).cpu().numpy()
    jh = np.einsum("tl,tqlc->tqc", j[hd.elem_dofs], hd.vec_shapes(bary))
    jex = np.cross(n, curlE / kappa)
    err = np.sqrt(np.sum(w * (np.abs(jh - jex) ** 2).sum(-1)) / np.sum(w * (np.abs(jex) ** 2).sum(-1)))
    print(f"Maxwell box, HCurl / HDivSurface 3: {hd.ndof} dofs, relative current error {err:.3e}", flush=True)
    assert err < 0.015
