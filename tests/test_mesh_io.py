"""Mesh readers: STL (ASCII and binary), Gmsh ASCII 2.2 / 4.1, Netgen .vol.

Round trips through files written here, and the ship surface that the
reference's Maxwell_Ship demo loads (demos/resources/ship_1.stl, kept as a
fixture under tests/golden): a closed, outward oriented surface on which the
double-layer potential of the density 1 is -1 inside (Gauss)."""
import os
import struct

import numpy as np

import ngbem_amd as nb
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def _tri_set(mesh):
    return sorted(tuple(np.round(mesh.xyz[t], 12).ravel()) for t in mesh.tri)


def _write_stl_ascii(mesh, path):
    with open(path, "w") as f:
        f.write("solid test\n")
        for t, n in zip(mesh.tri, mesh.normals()):
            f.write(f"facet normal {n[0]:.17g} {n[1]:.17g} {n[2]:.17g}\n outer loop\n")
            for v in t:
                f.write(f"  vertex {mesh.xyz[v][0]:.17g} {mesh.xyz[v][1]:.17g} {mesh.xyz[v][2]:.17g}\n")
            f.write(" endloop\nendfacet\n")
        f.write("endsolid test\n")


def _write_stl_binary(mesh, path):
    with open(path, "wb") as f:
        f.write(b"binary test".ljust(80, b" "))
        f.write(struct.pack("<I", mesh.nt))
        for t, n in zip(mesh.tri, mesh.normals()):
            f.write(struct.pack("<3f", *n))
            for v in t:
                f.write(struct.pack("<3f", *mesh.xyz[v]))
            f.write(b"\0\0")


def test_stl_round_trips(tmp_path):
    m = nb.ellipsoid(2, (1.0, 0.5, 2.0))
    a = tmp_path / "m.stl"
    _write_stl_ascii(m, a)
    r = nb.read_mesh(a)
    assert (r.nv, r.nt) == (m.nv, m.nt) and _tri_set(r) == _tri_set(m)
    assert np.allclose(np.sort(r.areas()), np.sort(m.areas()))
    b = tmp_path / "b.stl"
    _write_stl_binary(m, b)
    rb = nb.read_stl(b)
    assert (rb.nv, rb.nt) == (m.nv, m.nt)
    assert np.allclose(np.sort(rb.areas()), np.sort(m.areas()), rtol=1e-6)


def test_gmsh_and_netgen(tmp_path):
    m = nb.icosphere(1)
    g2 = tmp_path / "m2.msh"
    with open(g2, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % m.nv)
        for i, p in enumerate(m.xyz):
            f.write(f"{i + 1} {p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n")
        f.write("$EndNodes\n$Elements\n%d\n" % (m.nt + 1))
        f.write("1 15 2 7 1 1\n")  # a point element, skipped
        for i, t in enumerate(m.tri):
            f.write(f"{i + 2} 2 2 {3 + (i % 2)} 1 {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")
        f.write("$EndElements\n")
    r = nb.read_gmsh(g2)
    assert np.array_equal(r.tri, m.tri) and np.allclose(r.xyz, m.xyz)
    assert np.array_equal(r.regions, 3 + np.arange(m.nt) % 2)
    g4 = tmp_path / "m4.msh"
    with open(g4, "w") as f:
        f.write("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1 %d 1 %d\n2 5 0 %d\n" % (m.nv, m.nv, m.nv))
        for i in range(m.nv):
            f.write(f"{i + 1}\n")
        for p in m.xyz:
            f.write(f"{p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n")
        f.write("$EndNodes\n$Elements\n1 %d 1 %d\n2 5 2 %d\n" % (m.nt, m.nt, m.nt))
        for i, t in enumerate(m.tri):
            f.write(f"{i + 1} {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")
        f.write("$EndElements\n")
    r4 = nb.read_mesh(g4)
    assert np.array_equal(r4.tri, m.tri) and np.allclose(r4.xyz, m.xyz) and np.all(r4.regions == 5)
    v = tmp_path / "m.vol"
    with open(v, "w") as f:
        f.write("mesh3d\ndimension\n3\n\nsurfaceelements\n%d\n" % m.nt)
        for t in m.tri:
            f.write(f"1 2 1 0 3 {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")
        f.write("\npoints\n%d\n" % m.nv)
        for p in m.xyz:
            f.write(f"{p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n")
    rv = nb.read_mesh(v)
    assert np.array_equal(rv.tri, m.tri) and np.allclose(rv.xyz, m.xyz) and np.all(rv.regions == 2)


def test_orient_outward_flips_inverted_surfaces():
    m = nb.icosphere(1)
    inv = nb.Mesh(m.xyz, m.tri[:, [1, 0, 2]])
    assert np.allclose(nb.orient_outward(inv).normals(), m.normals())
    assert nb.orient_outward(m) is m


def test_ship_surface():
    ship = nb.read_stl(os.path.join(HERE, "golden", "ship_1.stl"))
    assert (ship.nv, ship.nt) == (347, 690)
    e = np.sort(np.concatenate([ship.tri[:, [0, 1]], ship.tri[:, [1, 2]], ship.tri[:, [2, 0]]]), axis=1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    assert np.all(counts == 2)  # closed
    assert nb.orient_outward(ship) is ship  # already outward
    # Gauss: double layer of 1 is -1 at interior points, 0 outside
    p0, p1, p2 = (ship.xyz[ship.tri[:, k]] for k in range(3))
    vol6 = np.einsum("ij,ij->i", p0, np.cross(p1, p2))
    centroid = ((p0 + p1 + p2) / 4.0 * vol6[:, None]).sum(0) / vol6.sum()
    far = centroid + np.array([0.0, 0.0, 10 * np.ptp(ship.xyz, axis=0).max()])
    dl = O.OracleOp(ship, O.LAPLACE_DL, 0.0, O.P0, O.P0, intorder=8, leafsize=10**6)
    u = dl.potential(np.ones(ship.nt), np.array([centroid, far]))[:, 0].real
    assert abs(u[0] + 1.0) < 1e-3 and abs(u[1]) < 1e-6


def test_box_mesh():
    m = nb.box(3, size=(2.0, 1.0, 0.5), centre=(0.1, 0.0, -0.2))
    assert m.nt == 6 * 2 * 9 and m.nv == 6 * 16 - 12 * 4 + 8
    p0, p1, p2 = (m.xyz[m.tri[:, k]] for k in range(3))
    assert np.isclose(np.einsum("ij,ij->i", p0, np.cross(p1, p2)).sum() / 6.0, 1.0)
    assert np.isclose(m.areas().sum(), 2 * (2.0 * 1.0 + 2.0 * 0.5 + 1.0 * 0.5))
    c = m.xyz[m.tri].mean(1) - np.array([0.1, 0.0, -0.2])
    assert np.all(np.einsum("ij,ij->i", c, m.normals()) > 0)
    e = np.sort(np.concatenate([m.tri[:, [0, 1]], m.tri[:, [1, 2]], m.tri[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert np.all(cnt == 2)
    assert m.GetBoundaries() == ("left", "right", "front", "back", "bottom", "top")
    assert m.Boundaries("top").mask.sum() == 18
