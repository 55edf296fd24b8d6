"""ctypes binding of libngbem_amd.so (C-ABI declared in include/ngbem_amd.h).

The library is built in-tree by ``ngbem_amd.build`` / ``__graft_entry__.build()``.
There is no CPU fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes
import os

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libngbem_amd.so")
_lib = None

SPACE_P0, SPACE_P1, SPACE_HDIV, SPACE_HCURL = 0, 1, 2, 3
(LAPLACE_SL, LAPLACE_DL, LAPLACE_HS, HELMHOLTZ_SL, HELMHOLTZ_DL, HELMHOLTZ_CF, HELMHOLTZ_HS, MAXWELL_SL,
 MAXWELL_DL, HELMHOLTZ_SL_DKAPPA) = range(10)
APPLY_TRANS, APPLY_CONJ, APPLY_MIXED, APPLY_ON_DEVICE, APPLY_PER_VECTOR, BATCH_EAGER, APPLY_GRAPH = 1, 2, 4, 8, 16, 32, 64
BATCH_CONSUME = 1
F64, C128, F32, C64 = 0, 1, 2, 3


class Params(ctypes.Structure):
    _fields_ = [("intorder", ctypes.c_int), ("leafsize", ctypes.c_int), ("eta", ctypes.c_double),
                ("eps", ctypes.c_double), ("dense", ctypes.c_int), ("mixed", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("nze", ctypes.c_int64), ("near_entries", ctypes.c_int64), ("far_factor_entries", ctypes.c_int64),
                ("n_blocks", ctypes.c_int32), ("n_near", ctypes.c_int32), ("n_far", ctypes.c_int32),
                ("n_fallback", ctypes.c_int32), ("max_rank", ctypes.c_int32), ("height", ctypes.c_int32),
                ("width", ctypes.c_int32), ("is_complex", ctypes.c_int32), ("n_singular_pairs", ctypes.c_int32 * 3),
                ("t_setup", ctypes.c_double), ("t_near", ctypes.c_double), ("t_far", ctypes.c_double),
                ("t_total", ctypes.c_double), ("near_kernel_evals", ctypes.c_double),
                ("far_kernel_evals", ctypes.c_double), ("t_kernel_near_regular", ctypes.c_double),
                ("t_kernel_near_singular", ctypes.c_double), ("t_kernel_aca", ctypes.c_double),
                ("t_kernel_near_gather", ctypes.c_double)]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if name == "n_singular_pairs" else v
        return d


class BlockInfo(ctypes.Structure):
    _fields_ = [("near", ctypes.c_int32), ("fallback", ctypes.c_int32), ("row_cluster", ctypes.c_int32),
                ("col_cluster", ctypes.c_int32), ("row0", ctypes.c_int32), ("m", ctypes.c_int32),
                ("col0", ctypes.c_int32), ("n", ctypes.c_int32), ("rank", ctypes.c_int32)]


# every symbol include/ngbem_amd.h declares
EXPORTS = ["bem_last_error", "bem_device_count", "bem_reserve", "bem_test_points", "bem_test_compression_aca", "bem_mesh_create", "bem_mesh_set_curved", "bem_mesh_destroy", "bem_space_create", "bem_space_create_on", "bem_space_set_dual_mapping",
           "bem_space_dofs",
           "bem_space_ndof", "bem_space_destroy", "bem_op_create", "bem_op_create_multi", "bem_op_assemble", "bem_op_apply",
           "bem_op_stats", "bem_op_autotune", "bem_op_profile_apply", "bem_probe_read", "bem_op_destroy", "bem_op_num_blocks", "bem_op_block_info", "bem_op_block_dofs",
           "bem_op_block_data", "bem_op_tree", "bem_op_potential_dim", "bem_op_potential", "bem_op_potential_noalloc", "bem_arnoldi_work_bytes", "bem_arnoldi_step", "bem_batch_create", "bem_batch_sizes", "bem_batch_apply",
           "bem_batch_destroy"]


def lib_path() -> str:
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"{_LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                           f"g.build()'` (no CPU fallback exists)")
    L = ctypes.CDLL(_LIB_PATH)
    P = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    L.bem_last_error.restype = ctypes.c_char_p
    L.bem_device_count.restype = ctypes.c_int
    L.bem_mesh_create.argtypes = [dp, ctypes.c_int, ip, ctypes.c_int, ctypes.POINTER(P)]
    L.bem_mesh_destroy.argtypes = [P]
    L.bem_space_create.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(P)]
    L.bem_space_create_on.argtypes = [P, ctypes.c_int, ctypes.c_int, P, ctypes.POINTER(P)]
    L.bem_space_set_dual_mapping.argtypes = [P, ctypes.c_int]
    L.bem_mesh_set_curved.argtypes = [P, ctypes.c_int, dp]
    L.bem_space_dofs.argtypes = [P, ip]
    L.bem_space_ndof.argtypes = [P]
    L.bem_space_destroy.argtypes = [P]
    L.bem_op_create.argtypes = [ctypes.c_int, ctypes.c_double, P, P, ctypes.POINTER(Params), ctypes.POINTER(P)]
    L.bem_op_create_multi.argtypes = [ctypes.c_int, dp, ctypes.c_int, ip, P, P, ctypes.POINTER(Params),
                                      ctypes.POINTER(P)]
    L.bem_op_assemble.argtypes = [P, P]
    L.bem_op_apply.argtypes = [P, ctypes.c_int, P, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
    L.bem_op_stats.argtypes = [P, ctypes.POINTER(Stats)]
    L.bem_op_profile_apply.argtypes = [P, ctypes.c_int, ctypes.c_int, dp, P]
    L.bem_probe_read.argtypes = [P, ctypes.c_int64, ctypes.c_int, P, ctypes.c_int, P]
    L.bem_op_destroy.argtypes = [P]
    L.bem_op_num_blocks.argtypes = [P]
    L.bem_op_block_info.argtypes = [P, ctypes.c_int, ctypes.POINTER(BlockInfo)]
    L.bem_op_block_dofs.argtypes = [P, ctypes.c_int, ip, ip]
    L.bem_op_block_data.argtypes = [P, ctypes.c_int, P]
    L.bem_op_tree.argtypes = [P, ctypes.c_int, ip, dp, dp, ctypes.c_int]
    L.bem_reserve.argtypes = [ctypes.c_double]
    L.bem_test_points.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint, dp, dp]
    L.bem_test_compression_aca.argtypes = [dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_double,
                                           ctypes.POINTER(ctypes.c_int), dp, dp]
    L.bem_op_potential_dim.argtypes = [P]
    L.bem_op_potential.argtypes = [P, P, P, P, ctypes.c_int, P, P, ctypes.c_int, P]
    L.bem_op_potential_noalloc.argtypes = [P, P, P, P, ctypes.c_int, P, P, ctypes.c_int, P]
    L.bem_op_autotune.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, P]
    L.bem_arnoldi_work_bytes.restype = ctypes.c_int64
    L.bem_arnoldi_work_bytes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.bem_arnoldi_step.argtypes = [P, ctypes.c_int64, ctypes.c_int, ctypes.c_int, P, P, P, P, ctypes.c_int64,
                                   ctypes.c_int, P]
    L.bem_batch_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.c_int, ctypes.POINTER(P)]
    L.bem_batch_sizes.argtypes = [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_int)]
    L.bem_batch_apply.argtypes = [P, ctypes.c_int, dp, P, dp, P, ctypes.c_int, P]
    L.bem_batch_destroy.argtypes = [P]
    _lib = L
    return L


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().bem_last_error().decode(errors="replace")
        raise RuntimeError(f"{what}: {msg}" if what else msg)


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
