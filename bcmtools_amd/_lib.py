"""ctypes binding of libbcm_hip.so (the C ABI declared in include/bcm_hip.h).

The shared library holds both the host octree builder and the HIP kernels.
Nothing here falls back to a CPU path: if the library is missing, every
entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbcm_hip.so")

# Status codes (include/bcm_hip.h).
OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_STATE, ERR_BALANCE, ERR_NAN, ERR_NOCONV, ERR_BUSY, ERR_IO = range(10)
STATUS_NAMES = {
    0: "OK", 1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_NOMEM", 4: "ERR_STATE", 5: "ERR_BALANCE",
    6: "ERR_NAN", 7: "ERR_NOCONV", 8: "ERR_BUSY", 9: "ERR_IO",
}

F64, F32 = 0, 1
X_M, X_P, Y_M, Y_P, Z_M, Z_P = range(6)
BC_INNER, BC_DIRICHLET, BC_NEUMANN, BC_PERIODIC = range(4)
DIV_FLAT, DIV_SIMPLE, DIV_SHELL, DIV_USER = range(4)
ORDER_Z, ORDER_HILBERT, ORDER_RANDOM = range(3)
NODE_BRANCH, NODE_LEAF_ACTIVE, NODE_LEAF_INACTIVE = range(3)
BCSET_ALL_DIRICHLET, BCSET_SOR1_X, BCSET_SOR1_Y, BCSET_SOR1_Z, BCSET_PER_FACE = range(5)
GF_FLAT, GF_F2C, GF_C2F, GF_ALL = 1, 2, 4, 7
GF_X, GF_Y, GF_Z = 8, 16, 32
GF_BLOCKWISE = 64
VC_AT_ONCE, VC_SEPARATE_XYZ, VC_SEPARATE_XYZ_SEPARATE_LEVEL_DIFF = 0, 1, 2
SOR_AUTO, SOR_FUSED_TASKS, SOR_FUSED_GATHER, SOR_HALO, SOR_COLOUR_TASKS, SOR_NUM_VARIANTS = range(6)
PREC_FP64, PREC_MIXED, PREC_FP32 = range(3)
SOLVE_ACTIVE, SOLVE_CONVERGED, SOLVE_NOCONV, SOLVE_FAILED = range(4)
(K_GHOST, K_DIRICHLET, K_SOR, K_JACOBI, K_SPMV, K_RESID, K_DOT, K_BLAS1, K_REDUCE,
 K_ASSEMBLE, K_SOR_HALO, K_JACOBI_HALO) = range(12)
MAX_SHELLS = 8


class BlockInfo(ctypes.Structure):
    _fields_ = [
        ("pedigree", ctypes.c_uint64),
        ("level", ctypes.c_int32),
        ("root_id", ctypes.c_int32),
        ("origin", ctypes.c_double * 3),
        ("block_size", ctypes.c_double * 3),
        ("nb_level_diff", ctypes.c_int32 * 6),
        ("nb_id", (ctypes.c_int32 * 4) * 6),
        ("nb_subface", ctypes.c_int32 * 6),
        ("bc_type", ctypes.c_int32 * 6),
        ("bc_id", ctypes.c_int32 * 6),
    ]


BLOCK_DTYPE = np.dtype(BlockInfo)


class TreeParams(ctypes.Structure):
    _fields_ = [
        ("root_n", ctypes.c_int32 * 3),
        ("periodic", ctypes.c_int32 * 3),
        ("root_origin", ctypes.c_double * 3),
        ("root_length", ctypes.c_double),
        ("divider", ctypes.c_int32),
        ("min_level", ctypes.c_int32),
        ("max_level", ctypes.c_int32),
        ("n_shells", ctypes.c_int32),
        ("shells", (ctypes.c_double * 5) * MAX_SHELLS),
        ("ordering", ctypes.c_int32),
        ("balance", ctypes.c_int32),
        ("bc_setter", ctypes.c_int32),
        ("shell_mode", ctypes.c_int32),
        ("shell_margin", ctypes.c_double),
        ("face_bc_type", ctypes.c_int32 * 6),
    ]


class KrylovParams(ctypes.Structure):
    _fields_ = [
        ("omega", ctypes.c_double),
        ("count_pre", ctypes.c_int32),
        ("count_max", ctypes.c_int32),
        ("epsilon", ctypes.c_double),
    ]


class IRParams(ctypes.Structure):
    _fields_ = [
        ("tol", ctypes.c_double),
        ("inner_tol", ctypes.c_double),
        ("inner_max", ctypes.c_int32),
        ("outer_max", ctypes.c_int32),
        ("count_pre", ctypes.c_int32),
        ("omega", ctypes.c_double),
        ("fp32_only", ctypes.c_int32),
        ("fp32_strict", ctypes.c_int32),
    ]


class SolveParams(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int32),
        ("tol", ctypes.c_double),
        ("max_iter", ctypes.c_int32),
        ("count_pre", ctypes.c_int32),
        ("omega", ctypes.c_double),
        ("refresh_max", ctypes.c_int32),
        ("refresh_delta", ctypes.c_double),
        ("check_every", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("restart_every", ctypes.c_int32),
        ("stall_window", ctypes.c_int32),
        ("stagnation", ctypes.c_int32),
    ]


class SolveStats(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("iters", ctypes.c_int32),
        ("refreshes", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("relres", ctypes.c_double),
    ]


class CommFace(ctypes.Structure):
    _fields_ = [
        ("block", ctypes.c_int32),
        ("face", ctypes.c_int32),
        ("subface", ctypes.c_int32),
        ("kind", ctypes.c_int32),
        ("offset", ctypes.c_int64),
    ]


COMM_SAME, COMM_F2C, COMM_C2F = 0, 1, 2
COMM_SEND, COMM_RECV = 0, 1
COMM_ANY_LEVEL = 2


class CommPeer(ctypes.Structure):
    _fields_ = [
        ("rank", ctypes.c_int32),
        ("nfaces", ctypes.c_int32),
        ("first", ctypes.c_int64),
        ("offset", ctypes.c_int64),
        ("bytes", ctypes.c_int64),
    ]


class ServerParams(ctypes.Structure):
    _fields_ = [
        ("hbm_budget", ctypes.c_double),
        ("max_active", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("max_iter", ctypes.c_int32),
        ("count_pre", ctypes.c_int32),
        ("omega", ctypes.c_double),
        ("check_every", ctypes.c_int32),
        ("batch_blocks", ctypes.c_int32),
        ("window_blocks", ctypes.c_int32),
        ("max_wait", ctypes.c_double),
    ]


class ServerResult(ctypes.Structure):
    _fields_ = [
        ("tag", ctypes.c_int64),
        ("status", ctypes.c_int32),
        ("iters", ctypes.c_int32),
        ("relres", ctypes.c_double),
        ("t_submit", ctypes.c_double),
        ("t_admit", ctypes.c_double),
        ("t_done", ctypes.c_double),
        ("nblocks", ctypes.c_int32),
        ("precision", ctypes.c_int32),
    ]


class BCMError(RuntimeError):
    def __init__(self, code, where):
        super().__init__(f"{where} failed: {STATUS_NAMES.get(code, code)}")
        self.code = code


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)

# Divider callback of bcm_tree_build_user: (user, pedigree) -> NODE_* kind.
DIVIDER_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64)

# name: (restype, argtypes)
SIGNATURES = {
    "bcm_tree_build": (_I, [ctypes.POINTER(TreeParams), ctypes.POINTER(_P)]),
    "bcm_tree_build_user": (_I, [ctypes.POINTER(TreeParams), DIVIDER_FN, _P, ctypes.POINTER(_P)]),
    "bcm_tree_set_threads": (_I, [_I]),
    "bcm_tree_num_blocks": (_I, [_P]),
    "bcm_tree_get_blocks": (_I, [_P, _P, _I]),
    "bcm_tree_destroy": (_I, [_P]),
    "bcm_mesh_create": (_I, [_P, _I, _I, _I, ctypes.POINTER(_P)]),
    "bcm_mesh_destroy": (_I, [_P]),
    "bcm_mesh_task_counts": (_I, [_P, _IP]),
    "bcm_mesh_sync": (_I, [_P]),
    "bcm_mesh_set_timing": (_I, [_P, _I]),
    "bcm_mesh_kernel_time": (_I, [_P, _I, _DP, ctypes.POINTER(ctypes.c_longlong)]),
    "bcm_mesh_reset_timing": (_I, [_P]),
    "bcm_field_create": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "bcm_field_destroy": (_I, [_P]),
    "bcm_host_alloc": (_I, [ctypes.c_size_t, ctypes.POINTER(_P)]),
    "bcm_host_free": (_I, [_P]),
    "bcm_field_upload": (_I, [_P, _P, ctypes.c_size_t]),
    "bcm_field_download": (_I, [_P, _P, ctypes.c_size_t]),
    "bcm_field_upload_blocks": (_I, [_P, _I, _I, _P, ctypes.c_size_t]),
    "bcm_field_download_blocks": (_I, [_P, _I, _I, _P, ctypes.c_size_t]),
    "bcm_field_dump2": (_I, [_P, ctypes.c_char_p, ctypes.c_char_p, _I, _I]),
    "bcm_field_load2": (_I, [_P, ctypes.c_char_p, ctypes.c_char_p, _I, _I]),
    "bcm_field_dump": (_I, [_P, ctypes.c_char_p, ctypes.c_char_p, _I, _I]),
    "bcm_field_load": (_I, [_P, ctypes.c_char_p, ctypes.c_char_p, _I, _I]),
    "bcm_field_write_vtk": (_I, [_P, ctypes.c_char_p, ctypes.c_char_p, _I, _I, _I, _I, _DP, _D]),
    "bcm_field_device_ptr": (_P, [_P]),
    "bcm_field_bytes": (ctypes.c_size_t, [_P]),
    "bcm_ghost_fill": (_I, [_P, _I]),
    "bcm_dirichlet_mirror": (_I, [_P, _DP, _I]),
    "bcm_impose_bc": (_I, [_P, _DP, _I]),
    "bcm_impose_face_bc": (_I, [_P, _IP, _DP]),
    "bcm_fold_face_bc": (_I, [ctypes.POINTER(_P), _P, _IP, _DP]),
    "bcm_rbsor_sweep": (_I, [_P, _P, _D]),
    "bcm_sor_iterate": (_I, [_P, _P, _D, _DP, _I, _I]),
    "bcm_sor_iterate_inner": (_I, [_P, _P, _D, _DP, _I, _I, _I]),
    "bcm_sor_residual": (_I, [_P, _P, _DP]),
    "bcm_ghost_fill_face": (_I, [_P, _I, _I, _I]),
    "bcm_comm_face_size": (ctypes.c_int64, [_P, _I]),
    "bcm_comm_list_create": (_I, [_P, ctypes.POINTER(CommFace), _I, ctypes.POINTER(_P)]),
    "bcm_comm_list_destroy": (_I, [_P]),
    "bcm_comm_pack_list": (_I, [_P, _P, _P, ctypes.c_size_t]),
    "bcm_comm_unpack_list": (_I, [_P, _P, _P, ctypes.c_size_t]),
    "bcm_comm_plan": (_I, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P, ctypes.c_int32,
                           ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(CommPeer), ctypes.c_int32,
                           ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(CommFace), ctypes.c_int64,
                           ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "bcm_comm_pack": (_I, [_P, ctypes.POINTER(CommFace), _I, _P, ctypes.c_size_t]),
    "bcm_comm_unpack": (_I, [_P, ctypes.POINTER(CommFace), _I, _P, ctypes.c_size_t]),
    "bcm_comm_pack_host": (_I, [_P, ctypes.POINTER(CommFace), _I, _P, ctypes.c_size_t]),
    "bcm_comm_unpack_host": (_I, [_P, ctypes.POINTER(CommFace), _I, _P, ctypes.c_size_t]),
    "bcm_vfield_create": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "bcm_vfield_destroy": (_I, [_P]),
    "bcm_vfield_upload": (_I, [_P, _P, ctypes.c_size_t]),
    "bcm_vfield_download": (_I, [_P, _P, ctypes.c_size_t]),
    "bcm_vfield_upload_blocks": (_I, [_P, _I, _I, _P, ctypes.c_size_t]),
    "bcm_vfield_download_blocks": (_I, [_P, _I, _I, _P, ctypes.c_size_t]),
    "bcm_vfield_ghost_fill": (_I, [_P, _I]),
    "bcm_vfield_ghost_fill_face": (_I, [_P, _I, _I, _I]),
    "bcm_sor_residual_noalloc": (_I, [_P, _P, _DP]),
    "bcm_sor_set_variant": (_I, [_P, _I]),
    "bcm_sor_get_variant": (_I, [_P]),
    "bcm_sor_autotune": (_I, [_P, _P, _D, _DP, _I, _I, _IP, _DP]),
    "bcm_jacobi7": (_I, [_P, _P, ctypes.POINTER(_P), _P, _D]),
    "bcm_calc_ax": (_I, [_P, ctypes.POINTER(_P), _P]),
    "bcm_calc_r": (_I, [_P, ctypes.POINTER(_P), _P, _P]),
    "bcm_calc_r2": (_I, [ctypes.POINTER(_P), _P, _P, _DP]),
    "bcm_dot": (_I, [_P, _P, _DP]),
    "bcm_axpy": (_I, [_P, _P, _D]),
    "bcm_xpay": (_I, [_P, _P, _D]),
    "bcm_axpyz": (_I, [_P, _P, _P, _D]),
    "bcm_axpbypz": (_I, [_P, _P, _P, _D, _D]),
    "bcm_fill": (_I, [_P, _D]),
    "bcm_copy": (_I, [_P, _P]),
    "bcm_jacobi7_mask": (_I, [_P, _P, ctypes.POINTER(_P), _P, _P, _D]),
    "bcm_calc_ax_mask": (_I, [_P, ctypes.POINTER(_P), _P, _P]),
    "bcm_calc_r_mask": (_I, [_P, ctypes.POINTER(_P), _P, _P, _P]),
    "bcm_calc_r2_mask": (_I, [ctypes.POINTER(_P), _P, _P, _P, _DP]),
    "bcm_dot_mask": (_I, [_P, _P, _P, _DP]),
    "bcm_axpy_mask": (_I, [_P, _P, _D, _P]),
    "bcm_xpay_mask": (_I, [_P, _P, _D, _P]),
    "bcm_axpyz_mask": (_I, [_P, _P, _P, _D, _P]),
    "bcm_axpbypz_mask": (_I, [_P, _P, _P, _D, _D, _P]),
    "bcm_copy_mask": (_I, [_P, _P, _P]),
    "bcm_setup_mask": (_I, [_P]),
    "bcm_update_virtual_cells": (_I, [_P, _I]),
    "bcm_impose_boundary_condition": (_I, [_P, _I, _DP, _I]),
    "bcm_jacobi_solve_mask": (_I, [_P, ctypes.POINTER(_P), _P, _P, _P, ctypes.POINTER(KrylovParams), _DP, _I, _IP,
                                   _DP]),
    "bcm_bicgstab_mask": (_I, [_P, ctypes.POINTER(_P), _P, ctypes.POINTER(_P), _P, ctypes.POINTER(KrylovParams), _DP,
                               _I, _IP, _DP]),
    "bcm_poisson_assemble": (_I, [ctypes.POINTER(_P), _P, _P, _DP, _I]),
    "bcm_jacobi7_poisson": (_I, [_P, _P, _P, _D]),
    "bcm_jacobi_poisson_iterate": (_I, [_P, _P, _D, _I]),
    "bcm_calc_ax_poisson": (_I, [_P, _P]),
    "bcm_calc_r_poisson": (_I, [_P, _P, _P]),
    "bcm_jacobi_solve": (_I, [_P, ctypes.POINTER(_P), _P, _P, ctypes.POINTER(KrylovParams), _DP, _I, _IP, _DP]),
    "bcm_bicgstab": (_I, [_P, ctypes.POINTER(_P), _P, ctypes.POINTER(_P), ctypes.POINTER(KrylovParams), _DP, _I,
                          _IP, _DP]),
    "bcm_solve_mixed": (_I, [_P, _P, _DP, _I, ctypes.POINTER(IRParams), _IP, _IP, _DP]),
    "bcm_solve_fp64": (_I, [_P, _P, _DP, _I, ctypes.POINTER(IRParams), _IP, _DP]),
    "bcm_mesh_create_batch": (_I, [_P, _IP, _I, _I, _I, ctypes.POINTER(_P)]),
    "bcm_mesh_num_segments": (_I, [_P]),
    "bcm_mesh_segment": (_I, [_P, _I, _IP, _IP]),
    "bcm_mesh_set_progress": (_I, [_P, _D]),
    "bcm_server_create": (_I, [ctypes.POINTER(ServerParams), ctypes.POINTER(_P)]),
    "bcm_server_destroy": (_I, [_P]),
    "bcm_server_clock": (_D, [_P]),
    "bcm_server_submit": (_I, [_P, _P, _I, _P, ctypes.c_uint64, _D, _I, _P, ctypes.c_int64]),
    "bcm_server_checkpoint": (_I, [_P, ctypes.c_int64, ctypes.c_char_p, ctypes.c_char_p, _I, _I, _IP]),
    "bcm_server_submit_restart": (_I, [_P, _P, _I, _P, ctypes.c_uint64, _D, _I, ctypes.c_char_p, ctypes.c_char_p,
                                       _I, _I, _P, ctypes.c_int64]),
    "bcm_server_progress": (_I, [_P, _D, _IP, _IP]),
    "bcm_server_poll": (_I, [_P, ctypes.POINTER(ServerResult), _I]),
    "bcm_server_stats": (_I, [_P, _DP, _IP, _IP]),
    "bcm_solve": (_I, [_P, _P, _DP, _I, ctypes.POINTER(SolveParams), _DP, ctypes.POINTER(SolveStats)]),
}

_lib = None


# The Fortran-convention per-block entry points (include/bcm_fortran.h):
# every argument is a pointer (device arrays; host scalars, sizes and
# results); none returns a value.  Name -> argument count.
FORTRAN_ARGC = {
    "fill_": 4, "fill_vf3d_": 5, "copy_": 4, "copy_integer_": 4, "add_": 5, "triad_": 6, "scal_": 4,
    "axpy_": 5, "xpay_": 5, "axpyz_": 6, "axpbypz_": 7, "dot_": 5, "dotx2_": 7,
    "jacobi_smoother_": 13, "jacobi_smoother2_": 7, "calc_ax_": 11, "calc_r_": 12, "calc_r2_": 12,
    "setup_mask_": 4, "copy_mask_": 5, "dot_mask_": 6, "calc_ax_mask_": 12, "jacobi_smoother_mask_": 14,
    "allreduce_": 2, "allreduce_max_": 2, "allreduce_min_": 2, "dot_all_": 5,
    "sup_copy_from_neighbor_": 7, "sup_copy_from_neighbor_c2f_": 7, "sup_copy_from_neighbor_f2c_": 7,
    "sup_copy_to_buffer_": 6, "sup_copy_to_buffer_c2f_": 6, "sup_copy_to_buffer_f2c_": 6,
    "sup_copy_from_buffer_": 6,
    "bc_x3_poiseuille_u_": 8, "bc_aw_poiseuille_u_": 10, "bc_x3_poiseuille_p_": 8, "bc_aw_poiseuille_p_": 10,
}
for _f in ("1", "2", "3", "4", "5", "6"):
    for _k in ("d", "d_f", "n"):
        FORTRAN_ARGC[f"bc_x{_f}_{_k}_"] = 4
for _f in ("1", "2", "5"):
    FORTRAN_ARGC[f"bc_x{_f}_p_"] = 4
for _f in ("w", "e", "s", "n", "b", "t"):
    for _k in ("d", "n"):
        FORTRAN_ARGC[f"bc_a{_f}_{_k}_"] = 6


def fortran(name):
    """A Fortran-convention entry point of libbcm_hip.so, typed."""
    fn = getattr(load(), name)
    fn.restype = None
    fn.argtypes = [_P] * FORTRAN_ARGC[name]
    return fn


def load():
    """Load libbcm_hip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP extension missing: {LIB_PATH} (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code, where):
    if code != OK:
        raise BCMError(code, where)
    return code


def dptr(arr):
    """ctypes double* of a C-contiguous float64 array (or None)."""
    if arr is None:
        return None
    return arr.ctypes.data_as(_DP)
