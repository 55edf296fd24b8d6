"""Flat triangle surface meshes: icospheres and ellipsoids.

A mesh is a pair of arrays: ``xyz`` (nv x 3, float64) and ``tri`` (nt x 3,
int32).  Triangles are oriented so that (p0-p2) x (p1-p2) points outward; that
is the normal NGSolve's surface element map F(xi,eta) = xi*p0 + eta*p1 +
(1-xi-eta)*p2 produces, and the one the reference's double-layer kernel and
common-edge rule assume (src/ngbem.cpp:583-586: neighbouring triangles run
through their shared edge in opposite directions).
"""
from __future__ import annotations

import re

import numpy as np


class Region:
    """A set of the mesh's surface elements, what NGSolve's
    ``mesh.Boundaries(pattern)`` returns (the reference takes it as
    ``definedon`` / ``trial_definedon`` / ``test_definedon``,
    src/python_bem.cpp:42-57): ``mask`` is one bool per triangle.  Regions
    combine with ``+`` (union), ``-`` (difference) and ``~`` (complement)."""

    def __init__(self, mesh, mask):
        self.mesh = mesh
        self.mask = np.ascontiguousarray(np.asarray(mask, dtype=bool).reshape(-1))
        if self.mask.shape != (mesh.nt,):
            raise ValueError("a region needs one flag per triangle")

    def _same(self, other):
        if not isinstance(other, Region) or other.mesh is not self.mesh:
            raise ValueError("regions of different meshes")

    def __add__(self, other):
        self._same(other)
        return Region(self.mesh, self.mask | other.mask)

    def __sub__(self, other):
        self._same(other)
        return Region(self.mesh, self.mask & ~other.mask)

    def __invert__(self):
        return Region(self.mesh, ~self.mask)

    def Mask(self):
        """One flag per region index (sorted region numbers), as NGSolve's
        Region.Mask() is one bit per boundary condition."""
        idx = np.unique(self.mesh.regions)
        return np.array([bool(self.mask[self.mesh.regions == r].all()) for r in idx])

    def elements(self):
        return np.flatnonzero(self.mask)

    def __repr__(self):
        return f"Region({int(self.mask.sum())} of {self.mesh.nt} elements)"


class Mesh:
    """Surface mesh made of flat triangles (vertex array + triangle array).

    ``regions``: a region number per triangle (default 0), ``region_names``:
    number -> name (unnamed numbers are called "default", as Netgen names an
    unnamed boundary)."""

    def __init__(self, xyz, tri, regions=None, region_names=None):
        self.xyz = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
        self.tri = np.ascontiguousarray(np.asarray(tri, dtype=np.int32).reshape(-1, 3))
        if self.tri.size and (self.tri.min() < 0 or self.tri.max() >= len(self.xyz)):
            raise ValueError("triangle vertex index out of range")
        self.regions = (np.zeros(self.nt, dtype=np.int32) if regions is None
                        else np.ascontiguousarray(np.asarray(regions, dtype=np.int32).reshape(-1)))
        if self.regions.shape != (self.nt,):
            raise ValueError("one region number per triangle")
        self.region_names = dict(region_names or {})
        self.geo_order = 1      # 1: flat triangles; 2-4 after Curve(order)
        self.nodes = None       # curved: nt x (k+1)(k+2)/2 x 3 Lagrange nodes
        self.project = None     # point -> nearest surface point (set by the generators)
        self._gpoly = None

    def Curve(self, order, project=None):
        """Curved (isoparametric) triangles of ``order`` (NGSolve's
        mesh.Curve, as the reference's tests mesh): each element's Lagrange
        nodes at (xi, eta) = (i / k, j / k) are the points of the flat triangle
        moved onto the surface by ``project`` (default: the generator's, e.g.
        the sphere for ``icosphere``; without one the nodes stay on the flat
        triangles).  Returns the mesh; call before creating spaces on it."""
        order = int(order)
        if getattr(self, "_bem_handle", None) is not None:
            raise RuntimeError("Curve must come before spaces are created on the mesh")
        if not 1 <= order <= 4:
            raise ValueError("geometry order 1 to 4")
        if order == 1:
            self.geo_order, self.nodes, self._gpoly = 1, None, None
            return self
        proj = project or self.project
        ref = lagrange_nodes(order)
        p0, p1, p2 = (self.xyz[self.tri[:, k]] for k in range(3))
        X = (ref[None, :, 0, None] * p0[:, None] + ref[None, :, 1, None] * p1[:, None]
             + (1.0 - ref[None, :, 0, None] - ref[None, :, 1, None]) * p2[:, None])
        if proj is not None:
            X = np.asarray(proj(X.reshape(-1, 3)), dtype=np.float64).reshape(X.shape)
        # the element's vertices are its corner nodes exactly
        k = order
        corner = {(k, 0): 0, (0, k): 1, (0, 0): 2}
        idx = 0
        for j in range(k + 1):
            for i in range(k + 1 - j):
                if (i, j) in corner:
                    X[:, idx] = self.xyz[self.tri[:, corner[(i, j)]]]
                idx += 1
        self.geo_order, self.nodes = order, np.ascontiguousarray(X)
        self._gpoly = np.einsum("mr,erc->emc", np.linalg.inv(_vandermonde(order, ref)), X)
        return self

    def geom(self, xi, eta):
        """Positions, unit normals and Jacobians at the reference points (xi,
        eta) of every element: (nt, q, 3), (nt, q, 3), (nt, q)."""
        return self.frames(xi, eta)[:3]

    def frames(self, xi, eta):
        """geom plus the tangents t1 = dX/dxi, t2 = dX/deta, (nt, q, 3) each."""
        xi, eta = np.atleast_1d(np.asarray(xi, float)), np.atleast_1d(np.asarray(eta, float))
        if self.geo_order == 1:
            p0, p1, p2 = (self.xyz[self.tri[:, k]] for k in range(3))
            t1, t2 = (p0 - p2)[:, None, :], (p1 - p2)[:, None, :]
            X = p2[:, None, :] + xi[None, :, None] * t1 + eta[None, :, None] * t2
            t1 = np.broadcast_to(t1, X.shape)
            t2 = np.broadcast_to(t2, X.shape)
        else:
            v, dx, dy = _monomials(self.geo_order, xi, eta)
            X = np.einsum("qm,emc->eqc", v, self._gpoly)
            t1 = np.einsum("qm,emc->eqc", dx, self._gpoly)
            t2 = np.einsum("qm,emc->eqc", dy, self._gpoly)
        c = np.cross(t1, t2)
        J = np.linalg.norm(c, axis=-1)
        return X, c / J[..., None], J, t1, t2

    def region_name(self, r) -> str:
        return self.region_names.get(int(r), "default")

    def GetBoundaries(self):
        """Names of the region numbers present, in increasing number order."""
        return tuple(self.region_name(r) for r in np.unique(self.regions))

    def Boundaries(self, pattern: str) -> Region:
        """The triangles whose region name matches the regular expression
        ``pattern`` as a whole (NGSolve's mesh.Boundaries)."""
        rx = re.compile(pattern)
        hit = [r for r in np.unique(self.regions) if rx.fullmatch(self.region_name(r))]
        return Region(self, np.isin(self.regions, hit))

    @property
    def nv(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def nt(self) -> int:
        return int(self.tri.shape[0])

    def edges(self):
        """Unique undirected edges (sorted vertex pairs), lexicographic order."""
        e = np.concatenate([self.tri[:, [0, 1]], self.tri[:, [1, 2]], self.tri[:, [2, 0]]])
        e = np.sort(e, axis=1)
        return np.unique(e, axis=0)

    def areas(self):
        if self.geo_order > 1:  # collapsed Gauss rule, exact for the polynomial part
            xi, eta, w = _collapsed_rule(self.geo_order + 3)
            return self.geom(xi, eta)[2] @ w
        p0, p1, p2 = (self.xyz[self.tri[:, k]] for k in range(3))
        return 0.5 * np.linalg.norm(np.cross(p0 - p2, p1 - p2), axis=1)

    def normals(self):
        """Unit normals (curved elements: at the centroid)."""
        if self.geo_order > 1:
            return self.geom(1.0 / 3.0, 1.0 / 3.0)[1][:, 0]
        p0, p1, p2 = (self.xyz[self.tri[:, k]] for k in range(3))
        c = np.cross(p0 - p2, p1 - p2)
        return c / np.linalg.norm(c, axis=1)[:, None]


def lagrange_nodes(k):
    """Reference coordinates (xi, eta) = (i / k, j / k), i + j <= k, j outer."""
    return np.array([(i / k, j / k) for j in range(k + 1) for i in range(k + 1 - j)], dtype=np.float64)


def _monomials(k, xi, eta):
    """xi^a eta^b (a + b <= k, degree outer, b inner) and their xi / eta
    derivatives at the points: (q, m) each."""
    v, dx, dy = [], [], []
    for d in range(k + 1):
        for b in range(d + 1):
            a = d - b
            v.append(xi ** a * eta ** b)
            dx.append(a * xi ** max(a - 1, 0) * eta ** b if a else np.zeros_like(xi))
            dy.append(b * xi ** a * eta ** max(b - 1, 0) if b else np.zeros_like(xi))
    return np.stack(v, 1), np.stack(dx, 1), np.stack(dy, 1)


def _vandermonde(k, ref):
    return _monomials(k, ref[:, 0], ref[:, 1])[0]


def _collapsed_rule(n):
    """Collapsed Gauss rule of n x n points on the reference triangle (weights sum to 1/2)."""
    g, gw = np.polynomial.legendre.leggauss(n)
    g, gw = 0.5 * (g + 1.0), 0.5 * gw
    U, V = np.meshgrid(g, g, indexing="ij")
    W = (gw[:, None] * gw[None, :]) * (1.0 - U)
    return U.ravel(), (V * (1.0 - U)).ravel(), W.ravel()


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1)[:, None]
    return v, f


def icosphere(level: int, radius: float = 1.0) -> Mesh:
    """Unit icosphere subdivided ``level`` times: 20*4^L triangles, 10*4^L+2 vertices.

    Each subdivision splits a triangle (a,b,c) into (a,ab,ca), (b,bc,ab),
    (c,ca,bc), (ab,bc,ca) and projects the new midpoints onto the sphere, so the
    numbering is deterministic and orientation is preserved.
    """
    v, f = _icosahedron()
    verts = [tuple(x) for x in v]
    for _ in range(level):
        mid = {}
        newf = []

        def midpoint(a, b):
            key = (a, b) if a < b else (b, a)
            k = mid.get(key)
            if k is None:
                p = (np.array(verts[a]) + np.array(verts[b])) * 0.5
                p /= np.linalg.norm(p)
                verts.append(tuple(p))
                k = len(verts) - 1
                mid[key] = k
            return k

        for a, b, c in f:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            newf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = np.array(newf, dtype=np.int64)
    xyz = np.array(verts, dtype=np.float64) * radius
    tri = np.asarray(f, dtype=np.int32)
    # orient outward: (p0-p2) x (p1-p2) . centroid > 0
    p0, p1, p2 = xyz[tri[:, 0]], xyz[tri[:, 1]], xyz[tri[:, 2]]
    s = np.einsum("ij,ij->i", np.cross(p0 - p2, p1 - p2), p0 + p1 + p2)
    flip = s < 0
    tri[flip] = tri[flip][:, [1, 0, 2]]
    m = Mesh(xyz, tri)
    m.project = lambda P: radius * P / np.linalg.norm(P, axis=-1, keepdims=True)
    return m


def ellipsoid(level: int, axes=(1.0, 1.0, 1.0)) -> Mesh:
    """Icosphere of ``level`` scaled by the semi-axes (orientation preserved);
    Curve projects radially in the sphere's coordinates."""
    m = icosphere(level)
    ax = np.asarray(axes, dtype=np.float64)
    e = Mesh(m.xyz * ax[None, :], m.tri)
    e.project = lambda P: ax * (P / ax) / np.linalg.norm(P / ax, axis=-1, keepdims=True)
    return e


def box(n, size=(1.0, 1.0, 1.0), centre=(0.0, 0.0, 0.0)) -> Mesh:
    """Surface of an axis-parallel box, each face a structured grid (n cells
    along every axis, or n = (nx, ny, nz) per axis) cut into two triangles per
    cell, oriented outward; faces are regions 1-6 named "left", "right",
    "front", "back", "bottom", "top" (-x, +x, -y, +y, -z, +z).  Flat triangles
    represent the box exactly."""
    ns = [int(n)] * 3 if np.isscalar(n) else [int(x) for x in n]
    if min(ns) < 1 or len(ns) != 3:
        raise ValueError("box needs n >= 1 (or three of them)")
    h = np.asarray(size, dtype=np.float64)
    c = np.asarray(centre, dtype=np.float64)
    ts = [np.linspace(-0.5, 0.5, k + 1) for k in ns]
    pts, tris, regions = [], [], []
    index = {}

    def vid(p):
        key = tuple(np.round(p, 12))
        if key not in index:
            index[key] = len(pts)
            pts.append(p)
        return index[key]

    # (axis, side): the face's outward normal is side * e_axis
    for r, (ax, side) in enumerate(((0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1))):
        u, v = [a for a in range(3) if a != ax]
        ids = np.zeros((ns[u] + 1, ns[v] + 1), dtype=np.int64)
        for i in range(ns[u] + 1):
            for j in range(ns[v] + 1):
                p = np.zeros(3)
                p[ax] = 0.5 * side
                p[u], p[v] = ts[u][i], ts[v][j]
                ids[i, j] = vid(p * h + c)
        for i in range(ns[u]):
            for j in range(ns[v]):
                a, b, cc, d = ids[i, j], ids[i + 1, j], ids[i + 1, j + 1], ids[i, j + 1]
                # (p0 - p2) x (p1 - p2) = e_u x e_v = sign * e_ax; flip for outward
                if np.cross(np.eye(3)[u], np.eye(3)[v])[ax] * side > 0:
                    tris += [[a, b, cc], [a, cc, d]]
                else:
                    tris += [[a, cc, b], [a, d, cc]]
                regions += [r + 1, r + 1]
    names = {1: "left", 2: "right", 3: "front", 4: "back", 5: "bottom", 6: "top"}
    return Mesh(np.array(pts), np.array(tris), regions, names)


# ------------------------------------------------------------ file formats --
# Surface triangle meshes from files.  The reference reads geometry through
# NGSolve/Netgen (e.g. the ship STL of demos/Maxwell_Ship.ipynb); here the
# common formats are read directly into flat-triangle meshes.  Vertex order is
# kept as stored (STL: counter-clockwise seen from the facet normal), so that
# (p0 - p2) x (p1 - p2) follows the file's orientation; ``orient_outward``
# flips a closed surface whose enclosed volume comes out negative.


def _merge_points(P, tri):
    """Merge coincident vertices (exact coordinate match)."""
    uniq, inv = np.unique(P, axis=0, return_inverse=True)
    return uniq, inv.reshape(-1)[tri]


def read_stl(path) -> Mesh:
    """ASCII or binary STL; coincident facet corners become one vertex."""
    data = open(path, "rb").read()
    is_ascii = data[:5].lower() == b"solid" and b"facet" in data[:2048]
    if not is_ascii:
        n = int(np.frombuffer(data[80:84], dtype="<u4")[0])
        if len(data) < 84 + 50 * n:
            raise ValueError("truncated binary STL")
        rec = np.frombuffer(data[84:84 + 50 * n], dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)),
                                                                   ("a", "<u2")]))
        P = rec["v"].reshape(-1, 3).astype(np.float64)
    else:
        vals = [line.split()[1:4] for line in data.decode("ascii", "replace").splitlines()
                if line.strip().lower().startswith("vertex")]
        P = np.array(vals, dtype=np.float64)
        if len(P) % 3:
            raise ValueError("ASCII STL: vertex count is not a multiple of 3")
    tri = np.arange(len(P)).reshape(-1, 3)
    xyz, tri = _merge_points(P, tri)
    return Mesh(xyz, tri)


def read_gmsh(path) -> Mesh:
    """Gmsh ASCII .msh, versions 2.2 and 4.1: the 3-node triangles (element type
    2); the physical tag becomes the element's region (v4: the surface
    entity's first physical tag from $Entities, else the entity tag), named
    by $PhysicalNames."""
    lines = open(path).read().split("\n")
    ver = 2.2
    nodes, tris, regions = {}, [], []
    names, surf_phys = {}, {}
    i = 0
    while i < len(lines):
        s = lines[i].strip()
        if s == "$MeshFormat":
            ver = float(lines[i + 1].split()[0])
        elif s == "$PhysicalNames":
            n = int(lines[i + 1])
            for k in range(n):
                f = lines[i + 2 + k].split(maxsplit=2)
                if int(f[0]) == 2:
                    names[int(f[1])] = f[2].strip().strip('"')
            i += n + 2
            continue
        elif s == "$Entities" and ver >= 4:
            # v4: elements carry entity tags; a surface entity's first
            # physical tag names its region
            npt, ncu, nsu, _ = (int(v) for v in lines[i + 1].split())
            j = i + 2 + npt + ncu
            for k in range(nsu):
                f = lines[j + k].split()
                nph = int(f[7])
                if nph > 0:
                    surf_phys[int(f[0])] = int(f[8])
            i = j + nsu
            continue
        elif s == "$Nodes":
            if ver < 4:
                n = int(lines[i + 1])
                for k in range(n):
                    f = lines[i + 2 + k].split()
                    nodes[int(f[0])] = [float(v) for v in f[1:4]]
                i += n + 2
            else:
                nb, _, _, _ = (int(v) for v in lines[i + 1].split())
                j = i + 2
                for _ in range(nb):
                    _, _, _, m = (int(v) for v in lines[j].split())
                    tags = [int(lines[j + 1 + k]) for k in range(m)]
                    for k in range(m):
                        nodes[tags[k]] = [float(v) for v in lines[j + 1 + m + k].split()[:3]]
                    j += 1 + 2 * m
                i = j
            continue
        elif s == "$Elements":
            if ver < 4:
                n = int(lines[i + 1])
                for k in range(n):
                    f = [int(v) for v in lines[i + 2 + k].split()]
                    if f[1] == 2:
                        nt = f[2]
                        tris.append(f[3 + nt:6 + nt])
                        regions.append(f[3] if nt > 0 else 0)
                i += n + 2
            else:
                nb, _, _, _ = (int(v) for v in lines[i + 1].split())
                j = i + 2
                for _ in range(nb):
                    _, tag, etype, m = (int(v) for v in lines[j].split())
                    for k in range(m):
                        f = [int(v) for v in lines[j + 1 + k].split()]
                        if etype == 2:
                            tris.append(f[1:4])
                            regions.append(surf_phys.get(tag, tag))
                    j += 1 + m
                i = j
            continue
        i += 1
    if not tris:
        raise ValueError("no triangles in the Gmsh file")
    keys = sorted(nodes)
    index = {k: n for n, k in enumerate(keys)}
    xyz = np.array([nodes[k] for k in keys])
    tri = np.array([[index[v] for v in t] for t in tris])
    return Mesh(xyz, tri, regions, names)


def read_netgen_vol(path) -> Mesh:
    """Netgen .vol: the "surfaceelements" (triangles; the boundary condition
    number becomes the region, named by the "bcnames" section) over the
    "points" section."""
    lines = [line.strip() for line in open(path).read().split("\n")]
    tris, regions, pts, names = [], [], [], {}
    i = 0
    while i < len(lines):
        s = lines[i]
        if s == "bcnames":
            n = int(lines[i + 1])
            for k in range(n):
                f = lines[i + 2 + k].split(maxsplit=1)
                if len(f) == 2:
                    names[int(f[0])] = f[1].strip()
            i += n + 2
            continue
        if s in ("surfaceelements", "surfaceelementsgi", "surfaceelementsuv"):
            n = int(lines[i + 1])
            for k in range(n):
                f = lines[i + 2 + k].split()
                npnt = int(f[4])
                if npnt == 3:
                    tris.append([int(v) - 1 for v in f[5:8]])
                    regions.append(int(f[1]))
            i += n + 2
            continue
        if s == "points":
            n = int(lines[i + 1])
            pts = [[float(v) for v in lines[i + 2 + k].split()[:3]] for k in range(n)]
            i += n + 2
            continue
        i += 1
    if not tris or not pts:
        raise ValueError("no surface triangles / points in the .vol file")
    return Mesh(np.array(pts), np.array(tris), regions, names)


def read_mesh(path) -> Mesh:
    """Dispatch on the file extension (.stl, .msh, .vol)."""
    p = str(path).lower()
    if p.endswith(".stl"):
        return read_stl(path)
    if p.endswith(".msh"):
        return read_gmsh(path)
    if p.endswith(".vol"):
        return read_netgen_vol(path)
    raise ValueError(f"unknown mesh format: {path}")


def orient_outward(mesh: Mesh) -> Mesh:
    """Flip every triangle of a closed, consistently oriented surface whose
    enclosed (signed) volume comes out negative."""
    p0, p1, p2 = (mesh.xyz[mesh.tri[:, k]] for k in range(3))
    vol = np.einsum("ij,ij->i", p0, np.cross(p1, p2)).sum() / 6.0
    if vol >= 0:
        return mesh
    return Mesh(mesh.xyz, mesh.tri[:, [1, 0, 2]], mesh.regions, mesh.region_names)
