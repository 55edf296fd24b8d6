"""C-ABI checks that need no GPU: the library loads, exports every function
include/ngbem_amd.h declares, rejects bad arguments with a message, and its
host-side cluster trees and block partitions are identical to the oracle's."""
import ctypes
import os
import re

import numpy as np
import pytest

import ngbem_amd as nb
from ngbem_amd import _capi
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "ngbem_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bem_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _capi.lib()
    names = header_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    assert sorted(names) == sorted(_capi.EXPORTS)


def test_bad_arguments_are_reported():
    L = _capi.lib()
    out = ctypes.c_void_p()
    xyz = np.zeros((3, 3))
    tri = np.array([[0, 1, 5]], dtype=np.int32)
    rc = L.bem_mesh_create(_capi.dptr(xyz), 3, _capi.iptr(tri), 1, ctypes.byref(out))
    assert rc == -1 and b"out of range" in L.bem_last_error()
    tri = np.array([[0, 1, 2]], dtype=np.int32)
    rc = L.bem_mesh_create(_capi.dptr(xyz), 3, _capi.iptr(tri), 1, ctypes.byref(out))
    assert rc == -1 and b"degenerate" in L.bem_last_error()
    m = nb.icosphere(1)
    with pytest.raises(RuntimeError, match="hypersingular operators need H1"):
        p0 = nb.SurfaceL2(m)
        params = _capi.Params(3, 40, 2.0, 1e-6, 0, 0)
        _capi.check(L.bem_op_create(_capi.LAPLACE_HS, 0.0, p0.handle, p0.handle, ctypes.byref(params),
                                    ctypes.byref(out)))
    with pytest.raises(NotImplementedError):
        nb.H1(m, order=5)
    with pytest.raises(NotImplementedError):
        nb.HDivSurface(m, order=5)
    with pytest.raises(NotImplementedError):
        nb.SurfaceL2(m, order=4)
    with pytest.raises(RuntimeError, match="hypersingular operators need H1"):
        l2 = nb.SurfaceL2(m, order=1)
        _capi.check(L.bem_op_create(_capi.LAPLACE_HS, 0.0, l2.handle, l2.handle, ctypes.byref(params),
                                    ctypes.byref(out)))
    with pytest.raises(TypeError):
        nb.MaxwellSingleLayerPotentialOperator(nb.SurfaceL2(m), 1.0)
    hd, hc = nb.HDivSurface(m), nb.HCurl(m)
    params = _capi.Params(3, 40, 2.0, 1e-6, 0, 0)
    for kern, tr, te, msg in ((_capi.MAXWELL_SL, p0, p0, "Maxwell single layer needs"),
                              (_capi.MAXWELL_SL, hd, hc, "Maxwell single layer needs"),
                              (_capi.MAXWELL_DL, hd, hc, "Maxwell double layer needs"),
                              (_capi.HELMHOLTZ_SL, hd, hd, "Laplace/Helmholtz kernels need")):
        with pytest.raises(RuntimeError, match=msg):
            _capi.check(L.bem_op_create(kern, 1.0, tr.handle, te.handle, ctypes.byref(params), ctypes.byref(out)))
    with pytest.raises(NotImplementedError):
        nb.HCurl(m, order=5)
    rc = L.bem_space_create(m._bem_handle.ptr, _capi.SPACE_HDIV, 5, ctypes.byref(out))
    assert rc == -1 and b"HDivSurface orders 0-4" in L.bem_last_error()
    rc = L.bem_space_create(m._bem_handle.ptr, _capi.SPACE_P1, 5, ctypes.byref(out))
    assert rc == -1 and b"H1 orders 1-4" in L.bem_last_error()
    # the Maxwell single layer takes one space kind on both sides
    hd1 = nb.HDivSurface(m, order=1)
    with pytest.raises(RuntimeError, match="Maxwell single layer needs"):
        _capi.check(L.bem_op_create(_capi.MAXWELL_SL, 1.0, hd1.handle, hd.handle, ctypes.byref(params),
                                    ctypes.byref(out)))


def op_create(space_trial, space_test, kernel, **p):
    L = _capi.lib()
    params = _capi.Params(p.get("intorder", 3), p.get("leafsize", 40), p.get("eta", 2.0), p.get("eps", 1e-6),
                          int(p.get("dense", False)), 0)
    out = ctypes.c_void_p()
    _capi.check(L.bem_op_create(kernel, p.get("kappa", 0.0), space_trial.handle, space_test.handle,
                                ctypes.byref(params), ctypes.byref(out)))
    return out


@pytest.mark.parametrize("level,trk,tek,leaf,eta", [(3, 0, 0, 40, 3.0), (4, 0, 0, 64, 2.0), (3, 1, 0, 20, 2.0),
                                                    (4, 1, 1, 40, 2.0), (3, 0, 1, 16, 3.0), (3, 3, 2, 32, 2.0),
                                                    (4, 2, 2, 64, 2.0)])
def test_tree_and_blocks_identical_to_oracle(level, trk, tek, leaf, eta):
    m = nb.icosphere(level)
    sp = {0: nb.SurfaceL2, 1: nb.H1, 2: nb.HDivSurface, 3: nb.HCurl}
    tr, te = sp[trk](m), sp[tek](m)
    kern = _capi.LAPLACE_DL if trk < 2 else (_capi.MAXWELL_DL if trk == 3 else _capi.MAXWELL_SL)
    h = op_create(tr, te, kern, leafsize=leaf, eta=eta, kappa=1.0)
    L = _capi.lib()
    try:
        ref = O.OracleOp(m, kern, kappa=1.0, trial=trk, test=tek, leafsize=leaf, eta=eta)
        for which, n in ((0, tr.ndof), (1, te.ndof)):
            perm = np.zeros(n, dtype=np.int32)
            ncl = L.bem_op_tree(h, which, _capi.iptr(perm), None, None, 0)
            cl = np.zeros(7 * ncl)
            cen = np.zeros(3 * ncl)
            L.bem_op_tree(h, which, _capi.iptr(perm), _capi.dptr(cl), _capi.dptr(cen), ncl)
            operm, ocl, ocen = ref.tree(which)
            assert np.array_equal(perm, operm)
            assert np.array_equal(cl.reshape(-1, 7), ocl)  # bit-identical radii too
            assert np.array_equal(cen.reshape(-1, 3), ocen)
        nblk = L.bem_op_num_blocks(h)
        assert nblk == ref.nblocks
        for b in range(nblk):
            bi = _capi.BlockInfo()
            L.bem_op_block_info(h, b, ctypes.byref(bi))
            rb = ref.block_info(b)
            assert (bi.row_cluster, bi.col_cluster) == (rb["row"], rb["col"])
            assert bool(bi.near and not bi.fallback) == rb["near"]
    finally:
        L.bem_op_destroy(h)


def test_dense_method_is_one_block():
    m = nb.icosphere(2)
    s = nb.SurfaceL2(m)
    h = op_create(s, s, _capi.LAPLACE_SL, dense=True)
    L = _capi.lib()
    assert L.bem_op_num_blocks(h) == 1
    bi = _capi.BlockInfo()
    L.bem_op_block_info(h, 0, ctypes.byref(bi))
    assert bi.near == 1 and bi.m == m.nt and bi.n == m.nt
    L.bem_op_destroy(h)


def _blocks_as_sets(h, offset=0):
    L = _capi.lib()
    out = set()
    for b in range(L.bem_op_num_blocks(h)):
        bi = _capi.BlockInfo()
        L.bem_op_block_info(h, b, ctypes.byref(bi))
        rows = np.zeros(bi.m, dtype=np.int32)
        cols = np.zeros(bi.n, dtype=np.int32)
        L.bem_op_block_dofs(h, b, _capi.iptr(rows), _capi.iptr(cols))
        out.add((tuple(sorted(rows - offset)), tuple(sorted(cols - offset)), bool(bi.near)))
    return out


@pytest.mark.parametrize("kind", [0, 1])
def test_multi_geometry_partition_is_per_geometry(kind):
    """Several geometries in one operator: no block couples two of them, and
    each geometry gets exactly the blocks of its own operator."""
    meshes = [nb.ellipsoid(2, (1.0, 0.7, 1.3)), nb.ellipsoid(3, (0.6, 1.1, 0.9)), nb.ellipsoid(2, (1.5, 1.0, 0.5))]
    U, eoff, voff = nb.union_mesh(meshes)
    sp = nb.SurfaceL2 if kind == 0 else nb.H1
    su = sp(U)
    L = _capi.lib()
    params = _capi.Params(3, 24, 2.0, 1e-6, 0, 0)
    kap = np.array([1.0, 2.0, 3.0])
    out = ctypes.c_void_p()
    _capi.check(L.bem_op_create_multi(_capi.HELMHOLTZ_SL, _capi.dptr(kap), 3, _capi.iptr(eoff.astype(np.int32)),
                                      su.handle, su.handle, ctypes.byref(params), ctypes.byref(out)))
    try:
        doff = eoff if kind == 0 else voff
        got = _blocks_as_sets(out)
        for i, m in enumerate(meshes):
            si = sp(m)
            h = op_create(si, si, _capi.HELMHOLTZ_SL, leafsize=24, kappa=kap[i])
            try:
                want = {(tuple(r + doff[i] for r in rows), tuple(c + doff[i] for c in cols), near)
                        for rows, cols, near in _blocks_as_sets(h)}
            finally:
                L.bem_op_destroy(h)
            mine = {b for b in got if doff[i] <= b[0][0] < doff[i + 1]}
            assert mine == want
            for rows, cols, _ in mine:
                assert doff[i] <= min(cols) and max(cols) < doff[i + 1]
        assert sum(1 for _ in got) == len(got)
    finally:
        L.bem_op_destroy(out)


def test_multi_geometry_rejects_bad_input():
    L = _capi.lib()
    m1 = nb.icosphere(1)
    U, eoff, _ = nb.union_mesh([m1, m1])
    su = nb.SurfaceL2(U)
    params = _capi.Params(3, 24, 2.0, 1e-6, 0, 0)
    out = ctypes.c_void_p()
    kap = np.array([1.0, 2.0])
    with pytest.raises(RuntimeError, match="element offsets"):
        bad = np.array([0, 10, 30], dtype=np.int32)
        _capi.check(L.bem_op_create_multi(_capi.HELMHOLTZ_SL, _capi.dptr(kap), 2, _capi.iptr(bad), su.handle,
                                          su.handle, ctypes.byref(params), ctypes.byref(out)))
    with pytest.raises(RuntimeError, match="share a vertex"):
        shared = nb.Mesh(U.xyz, np.where(U.tri >= m1.nv, U.tri - m1.nv, U.tri))  # second copy reuses the vertices
        ss = nb.SurfaceL2(shared)
        _capi.check(L.bem_op_create_multi(_capi.HELMHOLTZ_SL, _capi.dptr(kap), 2,
                                          _capi.iptr(eoff.astype(np.int32)), ss.handle, ss.handle,
                                          ctypes.byref(params), ctypes.byref(out)))
    hd = nb.HDivSurface(U)
    with pytest.raises(RuntimeError, match="Maxwell single layer"):
        _capi.check(L.bem_op_create_multi(_capi.MAXWELL_SL, _capi.dptr(kap), 2, _capi.iptr(eoff.astype(np.int32)),
                                          hd.handle, hd.handle, ctypes.byref(params), ctypes.byref(out)))
