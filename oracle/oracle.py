"""ctypes binding of the CPU FP64 oracle (oracle/bem_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libbem_oracle.so")
_lib = None

# kernel ids (same numbering as the product's C-ABI)
(LAPLACE_SL, LAPLACE_DL, LAPLACE_HS, HELMHOLTZ_SL, HELMHOLTZ_DL, HELMHOLTZ_CF, HELMHOLTZ_HS, MAXWELL_SL,
 MAXWELL_DL, HELMHOLTZ_SL_DKAPPA) = range(10)
# SurfaceL2 1-3; H1 2-4; edge orders 1-4
P0, P1, HDIV, HCURL, L2_1, L2_2, P2, L2_3, P3, HDIV1, HCURL1, HDIV2, HDIV3, HDIV4, HCURL2, HCURL3, HCURL4, P4 = \
    range(18)
EV_ID, EV_ROT, EV_HELM, EV_MAXW, EV_TR3 = range(5)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_mesh_new.restype = P
        L.orc_mesh_new.argtypes = [dp, ctypes.c_int, ip, ctypes.c_int]
        L.orc_mesh_free.argtypes = [P]
        L.orc_mesh_set_curved.argtypes = [P, ctypes.c_int, dp]
        L.orc_mesh_at.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_double, dp]
        L.orc_space_new.restype = P
        L.orc_space_new.argtypes = [P, ctypes.c_int]
        L.orc_space_free.argtypes = [P]
        L.orc_space_ndof.argtypes = [P]
        L.orc_space_dofs.argtypes = [P, ip]
        L.orc_space_eval.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     dp]
        L.orc_rule4.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int]
        L.orc_trig_rule.argtypes = [ctypes.c_int, dp, ctypes.c_int]
        L.orc_gauss01.argtypes = [ctypes.c_int, dp, dp]
        L.orc_op_new.restype = P
        L.orc_op_new.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_op_free.argtypes = [P]
        L.orc_op_nblocks.argtypes = [P]
        L.orc_op_block.argtypes = [P, ctypes.c_int, ip]
        L.orc_op_block_dofs.argtypes = [P, ctypes.c_int, ip, ip]
        L.orc_tree.argtypes = [P, ctypes.c_int, ip, dp, dp, ctypes.c_int]
        L.orc_calc_dense.argtypes = [P, ip, ctypes.c_int, ip, ctypes.c_int, dp]
        L.orc_calc_aca.argtypes = [P, ip, ctypes.c_int, ip, ctypes.c_int, dp, dp, ctypes.c_int]
        L.orc_assemble.restype = ctypes.c_double
        L.orc_assemble.argtypes = [P, ip, ctypes.c_int, ctypes.c_int]
        L.orc_get_block.argtypes = [P, ctypes.c_int, dp]
        L.orc_set_block.argtypes = [P, ctypes.c_int, ctypes.c_int, dp]
        L.orc_nze.restype = ctypes.c_long
        L.orc_nze.argtypes = [P]
        L.orc_potential.argtypes = [P, dp, dp, dp, ctypes.c_int, P, dp]
        L.orc_op_new_on.restype = P
        L.orc_op_new_on.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_op_sizes.argtypes = [P, ip, ip]
        L.orc_op_dofs.argtypes = [P, ctypes.c_int, ip]
        L.orc_test_points.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint, dp, dp]
        L.orc_test_compression_aca.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_uint, dp, dp, dp, ip, ip]
        L.orc_apply.restype = ctypes.c_double
        L.orc_apply.argtypes = [P, ctypes.c_int, dp, dp, ctypes.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def rule4(kind, order):
    q = lib().orc_rule4(kind, order, None, 0)
    out = np.zeros(5 * q)
    lib().orc_rule4(kind, order, _dp(out), q)
    return out.reshape(q, 5)


def trig_rule(order):
    q = lib().orc_trig_rule(order, None, 0)
    out = np.zeros(3 * q)
    lib().orc_trig_rule(order, _dp(out), q)
    return out.reshape(q, 3)


def gauss01(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().orc_gauss01(n, _dp(x), _dp(w))
    return x, w


def test_points(nx, ny, eta, seed=1):
    X, Y = np.zeros((nx, 3)), np.zeros((ny, 3))
    lib().orc_test_points(nx, ny, float(eta), seed, _dp(X), _dp(Y))
    return X, Y


def test_compression_aca(nx, ny, eta, eps, seed=1, factors=False):
    """(rank, Frobenius error) of TestCompressionACA restated; with factors=True
    also U (nx x r), V (r x ny) and the row / column pivots."""
    err = np.zeros(1)
    U, V = np.zeros((nx, nx)), np.zeros((nx, ny))
    pi, pj = np.zeros(nx, dtype=np.int32), np.zeros(nx, dtype=np.int32)
    r = lib().orc_test_compression_aca(nx, ny, float(eta), float(eps), seed, _dp(err), _dp(U), _dp(V), _ip(pi),
                                       _ip(pj))
    if not factors:
        return r, float(err[0])
    Uc = U.reshape(-1)[: r * nx].reshape(r, nx).T.copy()  # stored column by column
    Vr = V.reshape(-1)[: r * ny].reshape(r, ny).copy()
    return r, float(err[0]), Uc, Vr, pi[:r].copy(), pj[:r].copy()


class OracleMesh:
    def __init__(self, xyz, tri, geo_order=1, nodes=None):
        """geo_order > 1: curved elements through the Lagrange nodes (nt x
        (k+1)(k+2)/2 x 3, (xi, eta) = (i / k, j / k), j outer, i inner)."""
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.tri = np.ascontiguousarray(tri, dtype=np.int32)
        self.h = lib().orc_mesh_new(_dp(self.xyz), len(self.xyz), _ip(self.tri), len(self.tri))
        self.geo_order = int(geo_order)
        if self.geo_order > 1:
            self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
            if self.nodes.shape != (len(self.tri), (geo_order + 1) * (geo_order + 2) // 2, 3):
                raise ValueError("nodes: nt x (k+1)(k+2)/2 x 3")
            if lib().orc_mesh_set_curved(self.h, self.geo_order, _dp(self.nodes)) != 0:
                raise ValueError("geometry order 2 to 4")

    def at(self, e, xi, eta):
        """(position, unit normal, measure, surface gradients of lambda_0..2 (3 x 3))."""
        out = np.zeros(16)
        lib().orc_mesh_at(self.h, int(e), float(xi), float(eta), _dp(out))
        return out[:3], out[3:6], out[6], out[7:].reshape(3, 3)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_mesh_free(self.h)
            self.h = None


class OracleSpace:
    """A space of the oracle on a mesh: its dofs and shape functions."""

    def __init__(self, mesh, kind):
        if not isinstance(mesh, OracleMesh):
            mesh = OracleMesh(mesh.xyz, mesh.tri, getattr(mesh, "geo_order", 1), getattr(mesh, "nodes", None))
        self.mesh, self.kind = mesh, kind
        self.h = lib().orc_space_new(mesh.h, kind)
        self.ndof = lib().orc_space_ndof(self.h)
        out = np.zeros(30 * len(mesh.tri), dtype=np.int32)
        nloc = lib().orc_space_dofs(self.h, _ip(out))
        self.elem_dofs = out[: nloc * len(mesh.tri)].reshape(len(mesh.tri), nloc)

    def eval(self, ev, e, xi, eta, piola=False):
        """nloc x dim values of evaluator ev at (xi, eta) of element e."""
        out = np.zeros(120)
        n = lib().orc_space_eval(self.h, self.mesh.h, ev, int(e), float(xi), float(eta), int(piola), _dp(out))
        return out[:n].reshape(self.elem_dofs.shape[1], -1)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_space_free(self.h)
            self.h = None


class OracleOp:
    """Operator restated from the reference: tree + partition at construction;
    blocks are assembled on request (``assemble``)."""

    def __init__(self, mesh, kernel, kappa=0.0, trial=P0, test=P0, intorder=3, leafsize=40, eta=2.0, eps=1e-6,
                 trial_mask=None, test_mask=None):
        """trial_mask / test_mask: boolean per element: the space lives on
        those elements only (dofs numbered compactly), None: the whole mesh."""
        if not isinstance(mesh, OracleMesh):
            mesh = OracleMesh(mesh.xyz, mesh.tri, getattr(mesh, "geo_order", 1), getattr(mesh, "nodes", None))
        self.mesh = mesh
        self.kernel = kernel
        self._masks = [None if k is None else np.ascontiguousarray(k, dtype=np.uint8) for k in (trial_mask, test_mask)]
        for k in self._masks:
            if k is not None and k.shape != (len(mesh.tri),):
                raise ValueError("element mask needs one entry per triangle")
        mp = [None if k is None else k.ctypes.data_as(ctypes.c_void_p) for k in self._masks]
        self.h = lib().orc_op_new_on(mesh.h, kernel, float(kappa), trial, test, mp[0], mp[1], intorder, leafsize,
                                     float(eta), float(eps))
        if not self.h:
            raise ValueError("oracle: operator not created")
        a, b = ctypes.c_int(), ctypes.c_int()
        lib().orc_op_sizes(self.h, ctypes.byref(a), ctypes.byref(b))
        self.ntrial, self.ntest = a.value, b.value

    def elem_dofs(self, test=False):
        """Element -> dof table (nt x nloc), -1 outside the space's region."""
        out = np.zeros(30 * len(self.mesh.tri), dtype=np.int32)
        nloc = lib().orc_op_dofs(self.h, int(test), _ip(out))
        return out[: nloc * len(self.mesh.tri)].reshape(len(self.mesh.tri), nloc)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_op_free(self.h)
            self.h = None

    @property
    def nblocks(self):
        return lib().orc_op_nblocks(self.h)

    def block_info(self, b):
        info = np.zeros(7, dtype=np.int32)
        lib().orc_op_block(self.h, b, _ip(info))
        return dict(near=bool(info[0]), row=int(info[1]), col=int(info[2]), m=int(info[3]), n=int(info[4]),
                    rank=int(info[5]), fallback=bool(info[6]))

    def blocks(self):
        return [self.block_info(b) for b in range(self.nblocks)]

    def block_dofs(self, b):
        i = self.block_info(b)
        rows = np.zeros(i["m"], dtype=np.int32)
        cols = np.zeros(i["n"], dtype=np.int32)
        lib().orc_op_block_dofs(self.h, b, _ip(rows), _ip(cols))
        return rows, cols

    def tree(self, which=0):
        n = self.ntrial if which == 0 else self.ntest
        perm = np.zeros(n, dtype=np.int32)
        ncl = lib().orc_tree(self.h, which, _ip(perm), None, None, 0)
        cl = np.zeros(7 * ncl)
        cen = np.zeros(3 * ncl)
        lib().orc_tree(self.h, which, _ip(perm), _dp(cl), _dp(cen), ncl)
        return perm, cl.reshape(ncl, 7), cen.reshape(ncl, 3)

    def dense(self, trialdofs=None, testdofs=None):
        if trialdofs is None:
            trialdofs = np.arange(self.ntrial)
        if testdofs is None:
            testdofs = np.arange(self.ntest)
        c = np.ascontiguousarray(trialdofs, dtype=np.int32)
        r = np.ascontiguousarray(testdofs, dtype=np.int32)
        out = np.zeros((len(r), len(c)), dtype=np.complex128)
        lib().orc_calc_dense(self.h, _ip(c), len(c), _ip(r), len(r), _dp(out.view(np.float64)))
        return out

    def aca(self, trialdofs, testdofs):
        c = np.ascontiguousarray(trialdofs, dtype=np.int32)
        r = np.ascontiguousarray(testdofs, dtype=np.int32)
        R = lib().orc_calc_aca(self.h, _ip(c), len(c), _ip(r), len(r), None, None, 0)
        if R < 0:
            return None
        U = np.zeros((len(r), R), dtype=np.complex128)
        V = np.zeros((R, len(c)), dtype=np.complex128)
        lib().orc_calc_aca(self.h, _ip(c), len(c), _ip(r), len(r), _dp(U.view(np.float64)), _dp(V.view(np.float64)), R)
        return U, V

    def assemble(self, subset=None, nthreads=0):
        if subset is None:
            return lib().orc_assemble(self.h, None, 0, nthreads)
        s = np.ascontiguousarray(subset, dtype=np.int32)
        return lib().orc_assemble(self.h, _ip(s), len(s), nthreads)

    def get_block(self, b):
        i = self.block_info(b)
        m, n, r = i["m"], i["n"], i["rank"]
        if r == -2:
            raise RuntimeError("block not assembled")
        size = m * n if r == -1 else r * (m + n)
        out = np.zeros(size, dtype=np.complex128)
        lib().orc_get_block(self.h, b, _dp(out.view(np.float64)))
        if r == -1:
            return out.reshape(m, n)
        return out[: m * r].reshape(m, r), out[m * r:].reshape(r, n)

    def set_block(self, b, data, rank):
        d = np.ascontiguousarray(data, dtype=np.complex128)
        lib().orc_set_block(self.h, b, int(rank), _dp(d.view(np.float64)))

    def nze(self):
        return lib().orc_nze(self.h)

    def apply(self, x, trans=False, nthreads=0):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(self.ntrial if trans else self.ntest, dtype=np.complex128)
        t = lib().orc_apply(self.h, int(trans), _dp(x.view(np.float64)), _dp(y.view(np.float64)), nthreads)
        self.last_apply_seconds = t
        return y

    def potential(self, coef, points, normals=None, elmask=None):
        """PotentialCF values at points (npts x dim, complex); dim = trial evaluator dim;
        elmask: boolean per element, the elements that contribute (None: all)."""
        c = np.ascontiguousarray(coef, dtype=np.complex128)
        em = None if elmask is None else np.ascontiguousarray(elmask, dtype=np.uint8)
        P = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        N = None if normals is None else np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, 3)
        dim = 4
        out = np.zeros((len(P), dim), dtype=np.complex128)  # large enough for every evaluator
        d = lib().orc_potential(self.h, _dp(c.view(np.float64)), _dp(P), None if N is None else _dp(N), len(P),
                                None if em is None else em.ctypes.data_as(ctypes.c_void_p), _dp(out.view(np.float64)))
        return out.reshape(-1)[: len(P) * d].reshape(len(P), d)

    def densify(self):
        """Full matrix from the assembled blocks (small problems only)."""
        A = np.zeros((self.ntest, self.ntrial), dtype=np.complex128)
        for b in range(self.nblocks):
            rows, cols = self.block_dofs(b)
            blk = self.get_block(b)
            if isinstance(blk, tuple):
                blk = blk[0] @ blk[1]
            A[np.ix_(rows, cols)] += blk
        return A
