"""Spaces and operators on part of the mesh (definedon regions), on the CPU.

Pins of the oracle's masked operators:
- SurfaceL2 (one dof per element): the masked operator's entries are the
  whole-mesh entries of the region's elements, bit for bit (same element-pair
  integrals);
- every space: the operator with its trial (or test) space on a region A plus
  the one on the complement of A, both written into whole-mesh numbering,
  is the whole-mesh operator (each element's contribution lands in exactly
  one of the two), to rounding;
- the potential is linear in the element set: masked by A plus masked by
  the complement equals unmasked.
Then the product's host side: Boundaries() regex semantics, region names from
the mesh readers, compressed dof counts, element -> dof tables equal to the
oracle's, and cluster trees / block partitions on masked spaces bit-identical
to the oracle's.
"""
import ctypes

import numpy as np
import pytest

import ngbem_amd as nb
from ngbem_amd import _capi
from oracle import oracle as O


def capped_sphere(level, zcut=0.3):
    """Icosphere with region 1 "cap" (centroid z > zcut) and 2 "rest"."""
    m = nb.icosphere(level)
    c = m.xyz[m.tri].mean(axis=1)
    return nb.Mesh(m.xyz, m.tri, np.where(c[:, 2] > zcut, 1, 2), {1: "cap", 2: "rest"})


def _embed(D, rows_map, cols_map, shape):
    """Write D (masked numbering) into a whole-mesh matrix."""
    out = np.zeros(shape, dtype=D.dtype)
    out[np.ix_(rows_map, cols_map)] = D
    return out


def _dof_map(masked_tab, full_tab):
    """masked dof -> whole-mesh dof, from the two element -> dof tables."""
    n = masked_tab.max() + 1
    mp = np.full(n, -1, dtype=np.int64)
    on = masked_tab >= 0
    mp[masked_tab[on]] = full_tab[on]
    assert (mp >= 0).all()
    return mp


def test_boundaries_regex_and_algebra():
    m = capped_sphere(2)
    cap, rest = m.Boundaries("cap"), m.Boundaries("rest")
    assert cap.mask.sum() + rest.mask.sum() == m.nt and not (cap.mask & rest.mask).any()
    assert np.array_equal(m.Boundaries(".*").mask, np.ones(m.nt, bool))
    assert np.array_equal(m.Boundaries("cap|rest").mask, (cap + rest).mask)
    assert not m.Boundaries("ca").mask.any()  # whole-name match, as NGSolve
    assert np.array_equal((~cap).mask, rest.mask)
    assert not (cap - cap).mask.any()
    assert m.GetBoundaries() == ("cap", "rest")
    assert list(cap.Mask()) == [True, False]
    with pytest.raises(ValueError):
        cap + nb.icosphere(1).Boundaries(".*")


def test_reader_region_names(tmp_path):
    m = nb.icosphere(1)
    g2 = tmp_path / "m2.msh"
    with open(g2, "w") as f:
        f.write('$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$PhysicalNames\n2\n2 3 "inner"\n2 4 "outer"\n$EndPhysicalNames\n')
        f.write("$Nodes\n%d\n" % m.nv)
        for i, p in enumerate(m.xyz):
            f.write(f"{i + 1} {p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n")
        f.write("$EndNodes\n$Elements\n%d\n" % m.nt)
        for i, t in enumerate(m.tri):
            f.write(f"{i + 1} 2 2 {3 + (i % 2)} 1 {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")
        f.write("$EndElements\n")
    r = nb.read_gmsh(g2)
    assert r.GetBoundaries() == ("inner", "outer")
    assert np.array_equal(r.Boundaries("outer").mask, np.arange(m.nt) % 2 == 1)
    # v4.1: surface entity 5 carries physical tag 7 named "shell"
    g4 = tmp_path / "m4.msh"
    with open(g4, "w") as f:
        f.write('$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$PhysicalNames\n1\n2 7 "shell"\n$EndPhysicalNames\n')
        f.write("$Entities\n0 0 1 0\n5 -1 -1 -1 1 1 1 1 7 0\n$EndEntities\n")
        f.write("$Nodes\n1 %d 1 %d\n2 5 0 %d\n" % (m.nv, m.nv, m.nv))
        for i in range(m.nv):
            f.write(f"{i + 1}\n")
        for p in m.xyz:
            f.write(f"{p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n")
        f.write("$EndNodes\n$Elements\n1 %d 1 %d\n2 5 2 %d\n" % (m.nt, m.nt, m.nt))
        for i, t in enumerate(m.tri):
            f.write(f"{i + 1} {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")
        f.write("$EndElements\n")
    r4 = nb.read_mesh(g4)
    assert np.all(r4.regions == 7) and r4.GetBoundaries() == ("shell",)
    v = tmp_path / "m.vol"
    with open(v, "w") as f:
        f.write("mesh3d\ndimension\n3\n\nsurfaceelements\n%d\n" % m.nt)
        for i, t in enumerate(m.tri):
            f.write(f"1 {1 + i % 2} 1 0 3 {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")
        f.write("\npoints\n%d\n" % m.nv)
        for p in m.xyz:
            f.write(f"{p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n")
        f.write("\nbcnames\n2\n1 sphere\n2 screen\n")
    rv = nb.read_mesh(v)
    assert rv.GetBoundaries() == ("sphere", "screen")
    assert rv.Boundaries("screen").mask.sum() == m.nt // 2


@pytest.mark.parametrize("cls", [nb.SurfaceL2, nb.H1, nb.HDivSurface, nb.HCurl])
def test_space_on_region_counts_and_numbering(cls):
    m = capped_sphere(2)
    reg = m.Boundaries("cap")
    full, part = cls(m), cls(m, definedon=reg)
    tri = m.tri[reg.mask]
    if cls is nb.SurfaceL2:
        assert part.ndof == reg.mask.sum()
    elif cls is nb.H1:
        assert part.ndof == len(np.unique(tri))
    else:
        e = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
        assert part.ndof == len(np.unique(e, axis=0))
    assert (part.elem_dofs[~reg.mask] == -1).all()
    assert np.array_equal(part.elements, reg.elements())
    # compressed numbering keeps the whole-mesh order
    mp = _dof_map(part.elem_dofs, full.elem_dofs)
    assert np.all(np.diff(mp) > 0)
    # the oracle numbers the same way
    ref = O.OracleOp(m, O.LAPLACE_SL if cls in (nb.SurfaceL2, nb.H1) else O.MAXWELL_SL, 1.0,
                     trial=part.kind, test=part.kind, trial_mask=reg.mask, test_mask=reg.mask)
    assert ref.ntrial == part.ndof
    assert np.array_equal(ref.elem_dofs(), part.elem_dofs)
    # definedon by pattern, and the errors
    assert cls(m, definedon="cap").ndof == part.ndof
    with pytest.raises(ValueError):
        cls(m, definedon=nb.icosphere(2).Boundaries(".*"))
    with pytest.raises(ValueError):
        cls(m, definedon=m.Boundaries("nothing"))


def test_mass_matrix_on_regions():
    m = capped_sphere(2)
    cap = m.Boundaries("cap")
    for cls in (nb.SurfaceL2, nb.H1):
        full, part = cls(m), cls(m, definedon=cap)
        r, c, v = nb.spaces.mass_matrix_coo(part, part)
        M = np.zeros((part.ndof, part.ndof))
        np.add.at(M, (r, c), v)
        assert np.isclose(M.sum(), m.areas()[cap.mask].sum() * (1 if cls is nb.SurfaceL2 else 1))
        # H1 on the region: constant 1 integrates to the region's area row by row
        if cls is nb.H1:
            assert np.allclose(M @ np.ones(part.ndof), [m.areas()[np.any(m.tri == v_, 1) & cap.mask].sum() / 3
                                                         for v_ in _dof_map(part.elem_dofs, full.elem_dofs)])


def test_oracle_surfacel2_region_is_submatrix():
    m = capped_sphere(2)
    cap, rest = m.Boundaries("cap").mask, m.Boundaries("rest").mask
    for kern, kap in ((O.LAPLACE_SL, 0.0), (O.HELMHOLTZ_CF, 2.0)):
        full = O.OracleOp(m, kern, kap, intorder=3).dense()
        sub = O.OracleOp(m, kern, kap, trial_mask=cap, test_mask=rest, intorder=3)
        assert (sub.ntrial, sub.ntest) == (cap.sum(), rest.sum())
        assert np.array_equal(sub.dense(), full[np.ix_(np.flatnonzero(rest), np.flatnonzero(cap))])


@pytest.mark.parametrize("kern,trk,tek,side", [(O.LAPLACE_SL, O.P1, O.P1, "trial"), (O.LAPLACE_DL, O.P1, O.P0, "trial"),
                                               (O.LAPLACE_HS, O.P1, O.P1, "test"),
                                               (O.HELMHOLTZ_HS, O.P1, O.P1, "trial"),
                                               (O.MAXWELL_SL, O.HDIV, O.HDIV, "trial"),
                                               (O.MAXWELL_DL, O.HCURL, O.HDIV, "test")])
def test_oracle_region_and_complement_sum_to_whole(kern, trk, tek, side):
    m = capped_sphere(1, zcut=0.2)
    A = m.Boundaries("cap").mask
    kap = 1.5 if kern >= O.HELMHOLTZ_SL else 0.0
    full = O.OracleOp(m, kern, kap, trial=trk, test=tek, intorder=3)
    D = full.dense()
    acc = np.zeros_like(D)
    for mask in (A, ~A):
        kw = dict(trial_mask=mask) if side == "trial" else dict(test_mask=mask)
        part = O.OracleOp(m, kern, kap, trial=trk, test=tek, intorder=3, **kw)
        rmap = _dof_map(part.elem_dofs(test=True), full.elem_dofs(test=True))
        cmap = _dof_map(part.elem_dofs(), full.elem_dofs())
        acc += _embed(part.dense(), rmap, cmap, D.shape)
    assert np.abs(acc - D).max() <= 1e-13 * np.abs(D).max()


def test_oracle_potential_masks_are_linear():
    m = capped_sphere(2)
    A = m.Boundaries("cap").mask
    ref = O.OracleOp(m, O.HELMHOLTZ_DL, 2.0, trial=O.P1, test=O.P1, intorder=3)
    c = np.random.default_rng(1).standard_normal(ref.ntrial) + 0j
    pts = np.array([[0.1, 0.2, 2.5], [-3.0, 0.4, 0.1], [0.0, 0.0, 0.3]])
    u = ref.potential(c, pts)
    ua, ub = ref.potential(c, pts, elmask=A), ref.potential(c, pts, elmask=~A)
    assert np.abs(ua + ub - u).max() <= 1e-13 * np.abs(u).max()
    assert np.abs(ua).max() > 0 and np.abs(ub).max() > 0
    # a masked P0 space's potential = the whole space's with zeros outside
    p0 = O.OracleOp(m, O.LAPLACE_SL, 0.0, intorder=3)
    p0r = O.OracleOp(m, O.LAPLACE_SL, 0.0, intorder=3, trial_mask=A, test_mask=A)
    cr = np.random.default_rng(2).standard_normal(p0r.ntrial) + 0j
    cf = np.zeros(m.nt, complex)
    cf[np.flatnonzero(A)] = cr
    assert np.abs(p0r.potential(cr, pts) - p0.potential(cf, pts)).max() <= 1e-14 * np.abs(p0.potential(cf, pts)).max()


def _op_create(tr, te, kernel, **p):
    L = _capi.lib()
    params = _capi.Params(3, p.get("leafsize", 40), p.get("eta", 2.0), 1e-6, 0, 0)
    out = ctypes.c_void_p()
    _capi.check(L.bem_op_create(kernel, p.get("kappa", 0.0), tr.handle, te.handle, ctypes.byref(params),
                                ctypes.byref(out)))
    return out


@pytest.mark.parametrize("trk,tek,kern", [(0, 0, _capi.LAPLACE_SL), (1, 0, _capi.LAPLACE_DL), (1, 1, _capi.LAPLACE_HS),
                                          (2, 2, _capi.MAXWELL_SL), (3, 2, _capi.MAXWELL_DL)])
def test_tree_on_regions_identical_to_oracle(trk, tek, kern):
    m = capped_sphere(3, zcut=-0.2)
    A, B = m.Boundaries("cap"), m.Boundaries("rest") + m.Boundaries("cap")
    sp = {0: nb.SurfaceL2, 1: nb.H1, 2: nb.HDivSurface, 3: nb.HCurl}
    tr, te = sp[trk](m, definedon=A), sp[tek](m, definedon=B)
    h = _op_create(tr, te, kern, leafsize=24, kappa=1.0)
    L = _capi.lib()
    try:
        ref = O.OracleOp(m, kern, kappa=1.0, trial=trk, test=tek, leafsize=24, trial_mask=A.mask, test_mask=B.mask)
        for which, n in ((0, tr.ndof), (1, te.ndof)):
            perm = np.zeros(n, dtype=np.int32)
            ncl = L.bem_op_tree(h, which, _capi.iptr(perm), None, None, 0)
            cl, cen = np.zeros(7 * ncl), np.zeros(3 * ncl)
            L.bem_op_tree(h, which, _capi.iptr(perm), _capi.dptr(cl), _capi.dptr(cen), ncl)
            operm, ocl, ocen = ref.tree(which)
            assert np.array_equal(perm, operm) and np.array_equal(cl.reshape(-1, 7), ocl)
            assert np.array_equal(cen.reshape(-1, 3), ocen)
        assert L.bem_op_num_blocks(h) == ref.nblocks
        for b in range(ref.nblocks):
            bi = _capi.BlockInfo()
            L.bem_op_block_info(h, b, ctypes.byref(bi))
            rb = ref.block_info(b)
            assert (bi.row_cluster, bi.col_cluster) == (rb["row"], rb["col"])
            assert bool(bi.near and not bi.fallback) == rb["near"]
    finally:
        L.bem_op_destroy(h)


def test_multi_geometry_rejects_region_spaces():
    m = capped_sphere(1)
    s = nb.SurfaceL2(m, definedon=m.Boundaries("cap"))
    L = _capi.lib()
    params = _capi.Params(3, 40, 2.0, 1e-6, 0, 0)
    out = ctypes.c_void_p()
    kap = np.array([1.0])
    eoff = np.array([0, m.nt], dtype=np.int32)
    with pytest.raises(RuntimeError, match="part of the mesh"):
        _capi.check(L.bem_op_create_multi(_capi.HELMHOLTZ_SL, _capi.dptr(kap), 1, _capi.iptr(eoff), s.handle,
                                          s.handle, ctypes.byref(params), ctypes.byref(out)))
