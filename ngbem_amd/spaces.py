"""Boundary finite element spaces on flat triangle meshes.

``SurfaceL2(mesh, order=0)``: one dof per triangle, shape function 1
(NGSolve SurfaceL2, order 0).  ``H1(mesh, order=1)``: one dof per vertex,
barycentric hat functions (NGSolve H1, order 1, on a surface mesh).  These are
the spaces the reference's operators take (src/python_bem.cpp:27-173); dof
numbering: triangle index (SurfaceL2) and vertex index (H1).

``definedon=mesh.Boundaries(...)`` puts a space on part of the mesh: only the
region's triangles carry shape functions, and the dofs are the ones those
triangles touch, numbered compactly in the whole-mesh order (NGSolve's
definedon space wrapped in Compress, which is how the reference's demos build
them, demos/HelmholtzBurtonMiller.ipynb cell 4).  ``elem_dofs`` is the
element -> dof table, -1 outside the region.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi
from .mesh import Mesh, Region, _collapsed_rule


def _mesh_handle(mesh: Mesh):
    h = getattr(mesh, "_bem_handle", None)
    if h is None:
        L = _capi.lib()
        out = ctypes.c_void_p()
        _capi.check(L.bem_mesh_create(_capi.dptr(mesh.xyz), mesh.nv, _capi.iptr(mesh.tri), mesh.nt,
                                      ctypes.byref(out)), "bem_mesh_create")
        h = _Handle(out, L.bem_mesh_destroy)
        if getattr(mesh, "geo_order", 1) > 1:
            _capi.check(L.bem_mesh_set_curved(out, mesh.geo_order, _capi.dptr(mesh.nodes)), "bem_mesh_set_curved")
        mesh._bem_handle = h
    return h


class _Handle:
    def __init__(self, ptr, destroy):
        self.ptr = ptr
        self._destroy = destroy

    def __del__(self):
        if self.ptr:
            self._destroy(self.ptr)
            self.ptr = None


class FESpace:
    kind = -1        # C-ABI space kind (BEM_SPACE_*)
    orders = ()      # supported orders

    def __init__(self, mesh: Mesh, order: int, complex: bool = False, definedon=None, dirichlet=None,
                 dual_mapping: bool = False):
        if dual_mapping and not isinstance(self, SurfaceL2):
            raise NotImplementedError("dual_mapping is a SurfaceL2 option")
        if dual_mapping and getattr(mesh, "geo_order", 1) > 1:
            raise NotImplementedError("dual_mapping on curved meshes (its 1 / J varies over the element)")
        if order not in self.orders:
            raise NotImplementedError(f"{type(self).__name__} supports orders {list(self.orders)}")
        self.order = int(order)
        if isinstance(definedon, str):
            definedon = mesh.Boundaries(definedon)
        if definedon is not None:
            if not isinstance(definedon, Region) or definedon.mesh is not mesh:
                raise ValueError("definedon must be a Region of the space's mesh (mesh.Boundaries(...))")
            if not definedon.mask.any():
                raise ValueError("definedon region holds no element")
        self.mesh = mesh
        self.definedon = definedon
        self.is_complex = bool(complex)
        self.dirichlet = dirichlet
        self._mesh_h = _mesh_handle(mesh)
        L = _capi.lib()
        out = ctypes.c_void_p()
        mask = None
        if definedon is not None:
            self._mask = np.ascontiguousarray(definedon.mask, dtype=np.uint8)
            mask = self._mask.ctypes.data_as(ctypes.c_void_p)
        _capi.check(L.bem_space_create_on(self._mesh_h.ptr, self.kind, order, mask, ctypes.byref(out)),
                    "bem_space_create")
        self._h = _Handle(out, L.bem_space_destroy)
        self.ndof = int(L.bem_space_ndof(out))
        tab = np.zeros(30 * mesh.nt, dtype=np.int32)  # at most 30 local dofs (edge spaces of order 4)
        nloc = int(L.bem_space_dofs(out, _capi.iptr(tab)))
        if nloc < 1:
            _capi.check(nloc, "bem_space_dofs")
        self.elem_dofs = tab[: nloc * mesh.nt].reshape(mesh.nt, nloc).astype(np.int64)
        # NGSolve's dual_mapping: shapes divided by the Jacobian 2 |e| of their
        # triangle; dof_scale[d] = 1 / J of dof d's triangle (None: no scaling)
        self.dual_mapping = bool(dual_mapping)
        self.dof_scale = None
        if self.dual_mapping:
            _capi.check(L.bem_space_set_dual_mapping(out, 1), "bem_space_set_dual_mapping")
            sc = np.ones(self.ndof)
            on = self.elem_dofs[:, 0] >= 0
            J = 2.0 * mesh.areas()
            for l in range(nloc):
                sc[self.elem_dofs[on, l]] = 1.0 / J[on]
            self.dof_scale = sc

    @property
    def elements(self):
        """Triangles the space lives on."""
        return np.flatnonzero(self.elem_dofs[:, 0] >= 0)

    def FreeDofs(self):
        """NGSolve's ``fes.FreeDofs()``: a boolean array over the dofs, False for
        the dofs of the triangles in the ``dirichlet`` region (a regex over the
        region names, or a Region), True elsewhere."""
        free = np.ones(self.ndof, dtype=bool)
        if self.dirichlet is not None:
            reg = self.dirichlet if isinstance(self.dirichlet, Region) else self.mesh.Boundaries(self.dirichlet)
            d = self.elem_dofs[reg.mask]
            free[d[d >= 0]] = False
        return free

    @property
    def handle(self):
        return self._h.ptr

    def __repr__(self):
        return f"{type(self).__name__}(ndof={self.ndof}, order={self.order})"


def _jacobi(n, a, t):
    """Jacobi polynomial P_n^(a,0)(t) (three-term recurrence)."""
    p0, p1 = np.ones_like(t), 0.5 * ((a + 2.0) * t + a)
    if n == 0:
        return p0
    for k in range(2, n + 1):
        c = 2.0 * k + a
        p0, p1 = p1, ((c - 1.0) * (c * (c - 2.0) * t + a * a) * p1 - 2.0 * (k + a - 1.0) * (k - 1.0) * c * p0) / \
            (2.0 * k * (k + a) * (c - 2.0))
    return p1


def _dubiner(p, x, y):
    """Dubiner basis of degree <= p in x = lambda_0, y = lambda_1 (i outer, j
    inner): P_i((2x+y-1)/(1-y)) (1-y)^i P_j^(2i+1,0)(2y-1); rows = points."""
    w, s = 1.0 - y, 2.0 * x + y - 1.0
    out = []
    for i in range(p + 1):
        q = [np.ones_like(s), s]
        for k in range(2, i + 1):
            q.append(((2 * k - 1) * s * q[-1] - (k - 1) * w * w * q[-2]) / k)
        for j in range(p - i + 1):
            out.append(q[i] * _jacobi(j, 2.0 * i + 1.0, 2 * y - 1))
    return np.stack(out, axis=-1)


class SurfaceL2(FESpace):
    """SurfaceL2, orders 0-3: constants (dof = triangle), plus for orders 1-3
    the triangle's higher Dubiner functions (in the element's barycentrics
    x = lambda_0, y = lambda_1; dofs nt + e (nloc - 1) + k - 1).
    ``dual_mapping=True`` (NGSolve's flag): every shape divided by its
    triangle's Jacobian 2 |e| (``dof_scale``); ``shapes`` stays the reference
    values."""
    kind = _capi.SPACE_P0
    orders = (0, 1, 2, 3)

    def __init__(self, mesh, order=0, complex=False, definedon=None, dirichlet=None, dual_mapping=False):
        super().__init__(mesh, order, complex, definedon, dirichlet, dual_mapping)
        self.ikind = (0, 4, 5, 7)[self.order]

    def elem_signs(self):
        """Per element and local dof: +1 (SurfaceL2 functions are element-local)."""
        return np.ones(self.elem_dofs.shape)

    def shapes(self, bary):
        """Shape values at barycentric points (q x 3): q x nloc."""
        return _dubiner(self.order, bary[:, 0], bary[:, 1])


class H1(FESpace):
    """H1, orders 1-3 on the surface: vertex hats lambda_k; per edge
    (a, b) = (v_l, v_l+1) the bubble lambda_a lambda_b and, for order 3,
    s_l lambda_a lambda_b (lambda_b - lambda_a) with s_l = +1 when v_l < v_l+1;
    for order 3 the face bubble lambda_0 lambda_1 lambda_2.  Dofs: vertices,
    then p - 1 per edge (edges in sorted (min, max) order), then the faces."""
    kind = _capi.SPACE_P1
    orders = (1, 2, 3, 4)

    def __init__(self, mesh, order=1, complex=False, definedon=None, dirichlet=None):
        super().__init__(mesh, order, complex, definedon, dirichlet)
        self.ikind = (None, 1, 6, 8, 17)[self.order]

    def shapes(self, bary):
        """Shape values with every edge oriented +1 (multiply by elem_signs()):
        lambda_k; per edge (a, b) lambda_a lambda_b (lambda_b - lambda_a)^q,
        q < p - 1; faces B, B lambda_0, B lambda_1 (B = lambda_0 lambda_1
        lambda_2), the first (p-1)(p-2)/2."""
        l0, l1, l2 = bary[:, 0], bary[:, 1], bary[:, 2]
        v = [l0, l1, l2]
        lam = [l0, l1, l2]
        for e in range(3):
            a, b = lam[e], lam[(e + 1) % 3]
            for q in range(self.order - 1):
                v.append(a * b * (b - a) ** q)
        B = l0 * l1 * l2
        v += [B, B * l0, B * l1][: (self.order - 1) * (self.order - 2) // 2]
        return np.stack(v, axis=-1)

    def elem_signs(self):
        """Per element and local dof: the orientation factor of the odd edge
        functions (q = 1), +1 elsewhere."""
        sg = np.ones(self.elem_dofs.shape)
        if self.order >= 3:
            t = self.mesh.tri
            o = np.where(t < np.roll(t, -1, axis=1), 1.0, -1.0)
            pe = self.order - 1
            sg[:, [3 + pe * e + 1 for e in range(3)]] = o
        return sg


def _interior_index(p):
    """(c, u, v) of the interior functions lambda_(c+1) lambda_(c+2) q c_c of
    the order-p edge spaces: q = lambda_0^u lambda_1^v, u + v <= p - 2 (u
    outer) for c = 0, 1; q = lambda_0^u lambda_1^(p-2-u) for c = 2."""
    out = [(c, u, v) for c in (0, 1) for u in range(p - 1) for v in range(p - 1 - u)]
    return out + [(2, u, p - 2 - u) for u in range(p - 1)]


def _edge_fns(p, lam, sgn):
    """The order-p edge functions in the curls c_k = n x grad lambda_k at the
    barycentric points lam (q x 3), orientations sgn (nt x 3): weights W (nt,
    q, nloc, 3) with phi = sum_k W[..., k] c_k, D = J div phi (nt, q, nloc),
    and the flags of the order-0 functions (nloc,).  Per edge l = (a, b) =
    (l, l+1), t = lambda_b - lambda_a: function 0 = s (lambda_b c_a - lambda_a
    c_b) (J div = 2 s), function i >= 1 = n x grad(s^(i-1) lambda_a lambda_b
    t^(i-1)) (div 0); then the interior functions g c_c (_interior_index) with
    J div = dg/dlambda_(c+1) - dg/dlambda_(c+2)."""
    nt, Q, npe = sgn.shape[0], lam.shape[0], p + 1
    nloc = 3 if p == 0 else (p + 1) * (p + 2)
    W = np.zeros((nt, Q, nloc, 3))
    D = np.zeros((nt, Q, nloc))
    low = np.zeros(nloc, dtype=bool)
    for e in range(3):
        a, b = e, (e + 1) % 3
        s = sgn[:, e][:, None]
        t, ab = lam[:, b] - lam[:, a], lam[:, a] * lam[:, b]
        for i in range(npe):
            l = e * npe + i
            if i == 0:
                W[:, :, l, a] = s * lam[None, :, b]
                W[:, :, l, b] = -s * lam[None, :, a]
                D[:, :, l] = 2.0 * s
                low[l] = True
                continue
            q = i - 1
            sq = s if q % 2 else np.ones_like(s)
            tq, tq1 = t ** q, (t ** (q - 1) if q else np.zeros_like(t))
            W[:, :, l, a] = sq * (lam[:, b] * tq - q * ab * tq1)[None]
            W[:, :, l, b] = sq * (lam[:, a] * tq + q * ab * tq1)[None]
    for j, (c, u, v) in enumerate(_interior_index(p)):
        l = 3 * npe + j
        a, b = (c + 1) % 3, (c + 2) % 3
        qv = lam[:, 0] ** u * lam[:, 1] ** v
        dq = [u * lam[:, 0] ** max(u - 1, 0) * lam[:, 1] ** v if u else np.zeros(Q),
              v * lam[:, 0] ** u * lam[:, 1] ** max(v - 1, 0) if v else np.zeros(Q), np.zeros(Q)]
        ab = lam[:, a] * lam[:, b]
        W[:, :, l, c] = (ab * qv)[None]
        dg = [ab * dq[0], ab * dq[1], ab * dq[2]]
        dg[a] = dg[a] + lam[:, b] * qv
        dg[b] = dg[b] + lam[:, a] * qv
        D[:, :, l] = (dg[a] - dg[b])[None]
    return W, D, low


class _EdgeSpace(FESpace):
    """Edge elements of order p = 0-4 (all tangential fields of degree p).
    Order 0: dof = mesh edge (sorted (min, max) vertex pairs); local dof l of
    a triangle is its edge (v_l, v_{l+1}), signed +1 when v_l < v_{l+1}.
    Order p: per edge p + 1 functions (local l (p+1) + i, global i ne +
    edge), then (p+1)(p-1) interior functions (local 3 (p+1) + j, global
    (p+1) ne + e (p+1)(p-1) + j); the functions are those of _edge_fns."""

    orders = (0, 1, 2, 3, 4)

    def __init__(self, mesh, order=0, complex=False, definedon=None, dirichlet=None):
        super().__init__(mesh, order, complex, definedon, dirichlet)
        tri = mesh.tri
        npe = self.order + 1
        loc = np.stack([tri, np.roll(tri, -1, axis=1)], axis=2)  # nt x 3 x 2
        loc = np.repeat(loc, npe, axis=1)                        # nt x 3 (p+1) x 2
        lo, hi = loc.min(2), loc.max(2)
        self.edges = np.full((self.ndof, 2), -1, dtype=np.int32)  # dof -> sorted vertex pair (-1: interior)
        d = self.elem_dofs[:, : 3 * npe]
        on = d >= 0
        self.edges[d[on], 0] = lo[on]
        self.edges[d[on], 1] = hi[on]
        e3 = np.stack([tri, np.roll(tri, -1, axis=1)], axis=2)
        self.elem_sign = np.where(e3[:, :, 0] < e3[:, :, 1], 1.0, -1.0)  # nt x 3, per local edge

    def rt_shapes(self, pts_bary):
        """phi_l at barycentric points (q x 3) of every flat triangle: nt x q x
        3(l) x 3(xyz), phi_l = sign_l / J (x - p_opp(l)) (unit-flux
        Raviart-Thomas, the order-0 functions)."""
        m = self.mesh
        P = m.xyz[m.tri]  # nt x 3 x 3
        J = 2.0 * m.areas()
        X = np.einsum("qk,tkc->tqc", pts_bary, P)
        popp = P[:, [2, 0, 1], :]  # vertex opposite local edge l = vertex l+2
        s = self.elem_sign / J[:, None]
        return s[:, None, :, None] * (X[:, :, None, :] - popp[:, None, :, :])

    def frames(self, pts_bary):
        """Positions, unit normals, Jacobians and the curls c_k = n x
        grad_Gamma lambda_k = (t2, -t1, t1 - t2) / J at the points of every
        element (flat or curved): (nt, q, 3), (nt, q, 3), (nt, q), (nt, q, 3, 3)."""
        X, n, J, t1, t2 = self.mesh.frames(pts_bary[:, 0], pts_bary[:, 1])
        C = np.stack([t2, -t1, t1 - t2], axis=2) / J[..., None, None]
        return X, n, J, C

    def trace_shapes(self, pts_bary):
        """The tangential fields the Maxwell operators see, nt x q x nloc x 3:
        HDivSurface: the basis functions; HCurl: their rotated traces n x N
        (-phi for the order-0 functions, phi otherwise)."""
        _, _, _, C = self.frames(pts_bary)
        W, _, low = _edge_fns(self.order, pts_bary, self.elem_sign)
        t = np.einsum("tqlk,tqkc->tqlc", W, C)
        if isinstance(self, HCurl):
            t[:, :, low] *= -1.0
        return t

    def div_shapes(self, pts_bary):
        """The scalars of the Maxwell evaluators, nt x q x nloc: HDivSurface
        div phi, HCurl curl N (div phi for the order-0 functions, -div phi
        otherwise)."""
        _, _, J, _ = self.frames(pts_bary)
        _, D, low = _edge_fns(self.order, pts_bary, self.elem_sign)
        d = D / J[..., None]
        if isinstance(self, HCurl):
            d[:, :, ~low] *= -1.0
        return d

    def vec_shapes(self, pts_bary):
        """The basis functions themselves, nt x q x nloc x 3: HDivSurface phi;
        HCurl N = -n x (n x N)."""
        t = self.trace_shapes(pts_bary)
        if isinstance(self, HCurl):
            n = self.frames(pts_bary)[1][:, :, None, :]
            t = -np.cross(n, t)
        return t


class HDivSurface(_EdgeSpace):
    """HDivSurface, orders 0-4: unit-flux Raviart-Thomas (RWG-type) functions,
    the surface curls of the oriented edge functions lambda_a lambda_b t^q and
    the interior functions (all tangential fields of degree p)."""
    kind = _capi.SPACE_HDIV

    def __init__(self, mesh, order=0, complex=False, definedon=None, dirichlet=None):
        super().__init__(mesh, order, complex, definedon, dirichlet)
        self.ikind = (2, 9, 11, 12, 13)[self.order]


class HCurl(_EdgeSpace):
    """HCurl, orders 0-4, on the surface: Whitney edge functions N_l = n x
    phi_l, the gradients of the oriented edge functions, and interior
    functions whose rotated traces are those of HDivSurface."""
    kind = _capi.SPACE_HCURL

    def __init__(self, mesh, order=0, complex=False, definedon=None, dirichlet=None):
        super().__init__(mesh, order, complex, definedon, dirichlet)
        self.ikind = (3, 10, 14, 15, 16)[self.order]


def mass_matrix_coo(trial: FESpace, test: FESpace):
    """Entries of the surface mass matrix  int u . v ds  (test x trial), exact for
    the polynomial integrands on flat triangles (curved ones: Gauss rules of
    a few more points): rows, cols, values."""
    m = trial.mesh
    if isinstance(trial, _EdgeSpace) or isinstance(test, _EdgeSpace):
        if not (isinstance(trial, _EdgeSpace) and isinstance(test, _EdgeSpace)):
            raise NotImplementedError("mass matrix between an edge space and a scalar space")
        # collapsed Gauss rule exact for the polynomial integrands of degree 2p
        # (flat), with 3 more points a side for the curved elements' 1 / J
        n = max(trial.order, test.order) + 2 + (3 if getattr(m, "geo_order", 1) > 1 else 0)
        x, y, W = _collapsed_rule(n)
        bary = np.stack([x, y, 1.0 - x - y], axis=1)
        J = m.geom(x, y)[2]
        loc = np.einsum("q,tq,tqic,tqjc->tij", W, J, test.vec_shapes(bary), trial.vec_shapes(bary))
    elif getattr(m, "geo_order", 1) > 1:
        # curved elements: the Jacobian at each point of a collapsed Gauss rule
        # of 8 x 8 points (exact to degree 15 for the polynomial part)
        x, y, W = _collapsed_rule(8)
        bary = np.stack([x, y, 1.0 - x - y], axis=1)
        J = m.geom(x, y)[2]
        loc = np.einsum("q,tq,qi,qj->tij", W, J, test.shapes(bary), trial.shapes(bary))
        loc = loc * test.elem_signs()[:, :, None] * trial.elem_signs()[:, None, :]
    elif trial.order > 1 or test.order > 1 or isinstance(trial, SurfaceL2) and trial.order or \
            isinstance(test, SurfaceL2) and test.order:
        # higher orders: reference Gram matrix by a collapsed Gauss rule
        # (exact to degree 8), scaled by 2 * area
        g, gw = np.polynomial.legendre.leggauss(5)
        g, gw = 0.5 * (g + 1.0), 0.5 * gw
        U, Vv = np.meshgrid(g, g, indexing="ij")
        W = (gw[:, None] * gw[None, :]) * (1.0 - U)
        x, y = U.ravel(), (Vv * (1.0 - U)).ravel()
        bary = np.stack([x, y, 1.0 - x - y], axis=1)
        G = np.einsum("q,qi,qj->ij", W.ravel(), test.shapes(bary), trial.shapes(bary))
        loc = 2.0 * m.areas()[:, None, None] * G[None]
        loc = loc * test.elem_signs()[:, :, None] * trial.elem_signs()[:, None, :]
    else:
        area = m.areas()
        ni, nj = test.elem_dofs.shape[1], trial.elem_dofs.shape[1]
        if ni == 1 and nj == 1:
            loc = area[:, None, None]
        elif ni == 1 or nj == 1:  # SurfaceL2 x H1: area / 3 per vertex
            loc = np.broadcast_to((area / 3.0)[:, None, None], (m.nt, ni, nj))
        else:  # H1 x H1: area/12 * (1 + delta_ij)
            loc = area[:, None, None] * (np.array([[2, 1, 1], [1, 2, 1], [1, 1, 2]], dtype=np.float64) / 12.0)[None]
    ni, nj = loc.shape[1], loc.shape[2]
    rows = np.repeat(test.elem_dofs, nj, axis=1).reshape(-1)
    cols = np.tile(trial.elem_dofs, (1, ni)).reshape(-1)
    vals = np.ascontiguousarray(loc).reshape(-1)
    keep = (rows >= 0) & (cols >= 0)  # elements of both spaces' regions
    rows, cols, vals = rows[keep], cols[keep], vals[keep]
    # dual-mapped SurfaceL2: u = u_hat / J per triangle on either side
    if getattr(test, "dof_scale", None) is not None:
        vals = vals * test.dof_scale[rows]
    if getattr(trial, "dof_scale", None) is not None:
        vals = vals * trial.dof_scale[cols]
    return rows, cols, vals
