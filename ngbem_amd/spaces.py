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
        if isinstance(self, _EdgeSpace) and getattr(mesh, "geo_order", 1) > 1:
            raise NotImplementedError(f"{type(self).__name__} on curved meshes")
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
        tab = np.zeros(10 * mesh.nt, dtype=np.int32)
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
    orders = (1, 2, 3)

    def __init__(self, mesh, order=1, complex=False, definedon=None, dirichlet=None):
        super().__init__(mesh, order, complex, definedon, dirichlet)
        self.ikind = (None, 1, 6, 8)[self.order]

    def shapes(self, bary):
        """Shape values with every edge oriented +1 (multiply by elem_signs())."""
        l0, l1, l2 = bary[:, 0], bary[:, 1], bary[:, 2]
        v = [l0, l1, l2]
        lam = [l0, l1, l2]
        for e in range(3):
            a, b = lam[e], lam[(e + 1) % 3]
            if self.order >= 2:
                v.append(a * b)
            if self.order == 3:
                v.append(a * b * (b - a))
        if self.order == 3:
            v.append(l0 * l1 * l2)
        return np.stack(v, axis=-1)

    def elem_signs(self):
        """Per element and local dof: the orientation factor of the order-3 edge
        functions, +1 elsewhere."""
        sg = np.ones(self.elem_dofs.shape)
        if self.order == 3:
            t = self.mesh.tri
            o = np.where(t < np.roll(t, -1, axis=1), 1.0, -1.0)
            sg[:, [4, 6, 8]] = o
        return sg


class _EdgeSpace(FESpace):
    """Edge elements.  Order 0: dof = mesh edge (sorted (min, max) vertex
    pairs); local dof l of a triangle is its edge (v_l, v_{l+1}), signed +1 when
    v_l < v_{l+1}.  Order 1: local dofs 2l (the order-0 function of edge l) and
    2l+1 (the edge bubble's function, below); global dofs: the order-0 ones
    (edge index), then ne + edge index."""

    orders = (0, 1)

    def __init__(self, mesh, order=0, complex=False, definedon=None, dirichlet=None):
        super().__init__(mesh, order, complex, definedon, dirichlet)
        tri = mesh.tri
        loc = np.stack([tri, np.roll(tri, -1, axis=1)], axis=2)  # nt x 3 x 2
        if self.order == 1:
            loc = np.repeat(loc, 2, axis=1)                      # nt x 6 x 2
        lo, hi = loc.min(2), loc.max(2)
        self.edges = np.zeros((self.ndof, 2), dtype=np.int32)  # dof -> sorted vertex pair
        on = self.elem_dofs >= 0
        self.edges[self.elem_dofs[on], 0] = lo[on]
        self.edges[self.elem_dofs[on], 1] = hi[on]
        e3 = np.stack([tri, np.roll(tri, -1, axis=1)], axis=2)
        self.elem_sign = np.where(e3[:, :, 0] < e3[:, :, 1], 1.0, -1.0)  # nt x 3, per local edge

    def rt_shapes(self, pts_bary):
        """phi_l at barycentric points (q x 3) of every triangle: nt x q x 3(l) x 3(xyz),
        phi_l = sign_l / J (x - p_opp(l)) (unit-flux Raviart-Thomas)."""
        m = self.mesh
        P = m.xyz[m.tri]  # nt x 3 x 3
        J = 2.0 * m.areas()
        X = np.einsum("qk,tkc->tqc", pts_bary, P)
        popp = P[:, [2, 0, 1], :]  # vertex opposite local edge l = vertex l+2
        s = self.elem_sign / J[:, None]
        return s[:, None, :, None] * (X[:, :, None, :] - popp[:, None, :, :])

    def bubble_curls(self, pts_bary):
        """psi_l = n x grad(lambda_a lambda_b) for the edges (a, b) = (l, l+1):
        nt x q x 3(l) x 3(xyz)."""
        m = self.mesh
        P = m.xyz[m.tri]
        e1, e2 = P[:, 0] - P[:, 2], P[:, 1] - P[:, 2]
        G = np.stack([e1, e2], 1)                                  # nt x 2 x 3
        Ginv = np.linalg.inv(np.einsum("tic,tjc->tij", G, G))     # metric inverse
        g01 = np.einsum("tij,tjc->tic", Ginv, G)                   # grad lambda_0, lambda_1
        g = np.stack([g01[:, 0], g01[:, 1], -g01[:, 0] - g01[:, 1]], 1)  # nt x 3 x 3
        c = np.cross(m.normals()[:, None, :], g)                  # n x grad lambda_k
        lam = pts_bary
        out = []
        for l in range(3):
            a, b = l, (l + 1) % 3
            out.append(lam[None, :, a, None] * c[:, None, b, :] + lam[None, :, b, None] * c[:, None, a, :])
        return np.stack(out, axis=2)

    def trace_shapes(self, pts_bary):
        """The tangential fields the Maxwell operators see, nt x q x nloc x 3:
        HDivSurface: the basis functions; HCurl: their rotated traces n x N."""
        phi = self.rt_shapes(pts_bary)
        if isinstance(self, HCurl):
            phi = -phi                                    # n x (n x phi_l)
        if self.order == 0:
            return phi
        psi = self.bubble_curls(pts_bary)
        return np.stack([phi, psi], axis=3).reshape(phi.shape[0], phi.shape[1], 6, 3)

    def vec_shapes(self, pts_bary):
        """The basis functions themselves, nt x q x nloc x 3: HDivSurface phi_l
        (and psi_l); HCurl N_l = n x phi_l (and grad(lambda_a lambda_b) =
        -n x psi_l)."""
        t = self.trace_shapes(pts_bary)
        if isinstance(self, HCurl):
            n = self.mesh.normals()[:, None, None, :]
            t = -np.cross(n, t)                           # N = -n x (n x N)
        return t


class HDivSurface(_EdgeSpace):
    """HDivSurface, orders 0-1: unit-flux Raviart-Thomas (RWG-type) functions,
    for order 1 with the surface curls of the edge bubbles (all linear
    tangential fields)."""
    kind = _capi.SPACE_HDIV

    def __init__(self, mesh, order=0, complex=False, definedon=None, dirichlet=None):
        super().__init__(mesh, order, complex, definedon, dirichlet)
        self.ikind = (2, 9)[self.order]


class HCurl(_EdgeSpace):
    """HCurl, orders 0-1, on the surface: Whitney edge functions N_l = n x phi_l,
    for order 1 with the gradients of the edge bubbles."""
    kind = _capi.SPACE_HCURL

    def __init__(self, mesh, order=0, complex=False, definedon=None, dirichlet=None):
        super().__init__(mesh, order, complex, definedon, dirichlet)
        self.ikind = (3, 10)[self.order]


def mass_matrix_coo(trial: FESpace, test: FESpace):
    """Entries of the surface mass matrix  int u . v ds  (test x trial), exact for
    P0/P1 and the lowest-order edge spaces on flat triangles: rows, cols, values."""
    m = trial.mesh
    if getattr(m, "geo_order", 1) > 1:
        # curved elements: the Jacobian at each point of a collapsed Gauss rule
        # of 8 x 8 points (exact to degree 15 for the polynomial part)
        x, y, W = _collapsed_rule(8)
        bary = np.stack([x, y, 1.0 - x - y], axis=1)
        J = m.geom(x, y)[2]
        loc = np.einsum("q,tq,qi,qj->tij", W, J, test.shapes(bary), trial.shapes(bary))
        loc = loc * test.elem_signs()[:, :, None] * trial.elem_signs()[:, None, :]
    elif isinstance(trial, _EdgeSpace) or isinstance(test, _EdgeSpace):
        if not (isinstance(trial, _EdgeSpace) and isinstance(test, _EdgeSpace)):
            raise NotImplementedError("mass matrix between an edge space and a scalar space")
        # edge-midpoint rule, exact for the quadratic integrands (orders 0-1)
        bary = np.array([[0.5, 0.5, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5]])
        w = m.areas()[:, None] / 3.0
        phi_i = test.vec_shapes(bary)
        phi_j = trial.vec_shapes(bary)
        loc = np.einsum("tq,tqic,tqjc->tij", w, phi_i, phi_j)  # nt x nloc x nloc
    elif trial.order > 1 or test.order > 1 or isinstance(trial, SurfaceL2) and trial.order or \
            isinstance(test, SurfaceL2) and test.order:
        # higher orders: reference Gram matrix by a collapsed Gauss rule
        # (exact to degree 7), scaled by 2 * area
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
