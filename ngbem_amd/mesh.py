"""Flat triangle surface meshes: icospheres and ellipsoids.

A mesh is a pair of arrays: ``xyz`` (nv x 3, float64) and ``tri`` (nt x 3,
int32).  Triangles are oriented so that (p0-p2) x (p1-p2) points outward; that
is the normal NGSolve's surface element map F(xi,eta) = xi*p0 + eta*p1 +
(1-xi-eta)*p2 produces, and the one the reference's double-layer kernel and
common-edge rule assume (src/ngbem.cpp:583-586: neighbouring triangles run
through their shared edge in opposite directions).
"""
from __future__ import annotations

import numpy as np


class Mesh:
    """Surface mesh made of flat triangles (vertex array + triangle array)."""

    def __init__(self, xyz, tri):
        self.xyz = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
        self.tri = np.ascontiguousarray(np.asarray(tri, dtype=np.int32).reshape(-1, 3))
        if self.tri.size and (self.tri.min() < 0 or self.tri.max() >= len(self.xyz)):
            raise ValueError("triangle vertex index out of range")

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
        p0, p1, p2 = (self.xyz[self.tri[:, k]] for k in range(3))
        return 0.5 * np.linalg.norm(np.cross(p0 - p2, p1 - p2), axis=1)

    def normals(self):
        p0, p1, p2 = (self.xyz[self.tri[:, k]] for k in range(3))
        c = np.cross(p0 - p2, p1 - p2)
        return c / np.linalg.norm(c, axis=1)[:, None]


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
    return Mesh(xyz, tri)


def ellipsoid(level: int, axes=(1.0, 1.0, 1.0)) -> Mesh:
    """Icosphere of ``level`` scaled by the semi-axes (orientation preserved)."""
    m = icosphere(level)
    return Mesh(m.xyz * np.asarray(axes, dtype=np.float64)[None, :], m.tri)
