"""ctypes binding of libcomech_hip.so (declared in include/comech_fe.h).

There is no CPU fallback: if the library is missing or fails to load, importing the solver raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcomech_hip.so")

c_int, c_ll, c_double, c_void_p = ctypes.c_int, ctypes.c_longlong, ctypes.c_double, ctypes.c_void_p
P_int, P_ll, P_double, P_u8 = (ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong),
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint8))


class SolveInfo(ctypes.Structure):
    _fields_ = [("iters", c_int), ("flag", c_int), ("outer", c_int), ("levels", c_int), ("relres", c_double),
                ("t_solve_ms", c_double), ("t_setup_ms", c_double)]


# name -> argtypes (all return int status unless listed in _RET)
SIGNATURES = {
    "cm_last_error": [],
    "cm_version": [],
    "cm_device_count": [P_int],
    "cm_sys_create": [ctypes.POINTER(c_void_p), c_void_p, c_int, c_int, c_int, P_double, P_double, P_double, c_int],
    "cm_sys_destroy": [c_void_p],
    "cm_sys_set_labels": [c_void_p, c_int, P_double, P_double, P_u8],
    "cm_sys_set_cellmat": [c_void_p, P_double, P_double, P_u8],
    "cm_sys_set_split": [c_void_p, c_ll, P_ll, c_ll, P_ll, P_ll],
    "cm_sys_set_extra": [c_void_p, c_ll, P_ll, P_ll, P_double],
    "cm_sys_set_rank1": [c_void_p, c_ll, P_ll, P_ll, P_double, P_double],
    "cm_sys_set_used": [c_void_p, c_ll, P_u8],
    "cm_sys_set_rigid": [c_void_p, c_int, c_ll, P_ll, P_ll, P_double, P_double],
    "cm_sys_build": [c_void_p, c_int, c_int, c_int, c_int, c_int, c_double, c_int, c_int],
    "cm_sys_levels": [c_void_p, P_int, P_ll, c_int],
    "cm_sys_level_stats": [c_void_p, c_int, P_ll, c_int],
    "cm_sys_apply": [c_void_p, c_int, c_int, c_void_p, c_void_p],
    "cm_sys_vcycle": [c_void_p, c_int, c_void_p, c_void_p],
    "cm_sys_level_data": [c_void_p, c_int, c_int, c_int, c_void_p],
    "cm_sys_set_fallback": [c_void_p, c_int, c_double],
    "cm_sys_power": [c_void_p, c_int, c_int, c_void_p, P_double],
    "cm_sys_smooth": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int],
    "cm_sys_solve": [c_void_p, c_int, c_void_p, c_void_p, c_double, c_int, c_int, c_double, ctypes.POINTER(SolveInfo)],
    "cm_sys_prepare_solve": [c_void_p],
    "cm_batch_create": [ctypes.POINTER(c_void_p), c_void_p, c_int],
    "cm_batch_destroy": [c_void_p],
    "cm_batch_add": [c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_int, P_int],
    "cm_batch_step": [c_void_p, c_int, P_int],
    "cm_batch_poll": [c_void_p, P_int, ctypes.POINTER(SolveInfo)],
    "cm_cell_stress": [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_ll, c_void_p, c_void_p, c_void_p,
                       c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "cm_cell_stress_ortho": [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_ll, c_void_p, c_void_p,
                             c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "cm_element_forces": [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_ll, c_void_p,
                          c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "cm_sys_time_apply": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_int, P_double],
    "cm_sys_autotune": [c_void_p, c_int, c_int],
    "cm_sys_set_tuned_kz": [c_void_p, c_int, c_int, c_int, c_int],
    "cm_sys_set_march_blocks": [c_void_p, c_int],
    "cm_sys_timings": [c_void_p, P_double, c_int],
    "cm_sys_set_kz": [c_void_p, c_int],
    "cm_sys_set_block_skip": [c_void_p, c_int],
    "cm_sys_set_copy_blocks": [c_void_p, c_int],
    "cm_sys_set_last_coarsening": [c_void_p, c_int],
    "cm_sys_set_graphs": [c_void_p, c_int],
    "cm_lattice_materials": [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_int,
                             c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "cm_bank_pack": [c_void_p, c_ll, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                     c_void_p, c_void_p],
    "cm_part_nodes": [c_void_p, c_ll, c_void_p, c_void_p, c_void_p, c_ll, c_void_p, c_void_p],
    "cm_body_fractions": [c_void_p, c_ll, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p],
    "cm_nearest_bone": [c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p,
                        c_void_p, c_void_p],
    "cm_spring_patches": [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p,
                          c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "cm_split_connectivity": [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                              c_ll, P_ll, P_ll],
    "cm_sys_sync": [c_void_p],
    "cm_sample_fe": [c_void_p] + [c_int] * 7 + [c_void_p] * 9 + [c_void_p] * 5 + [c_void_p] * 4,
    "cm_sample_nodes": [c_void_p] + [c_int] * 3 + [c_void_p] * 6 + [c_int] + [c_void_p] * 2 + [c_int] * 3 +
                       [c_void_p] * 9,
    "cm_sample_nodes_noalloc": [c_void_p] + [c_int] * 3 + [c_void_p] * 6 + [c_int] + [c_void_p] * 2 + [c_int] * 3 +
                               [c_void_p] * 7,
}
# include/comech_learned.h (the learned warm start's kernels, same library)
c_float = ctypes.c_float
SIGNATURES_NN = {
    "cm_nn_last_error": [],
    "cm_nn_gemm_tn": [c_void_p, c_int, c_int, c_int, c_void_p, c_ll, c_void_p, c_ll, c_void_p, c_ll, c_void_p, c_void_p,
                      c_ll, c_int, c_int, c_int, c_float],
    "cm_nn_im2col": [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_ll],
    "cm_nn_transpose": [c_void_p, c_ll, c_ll, c_int, c_void_p, c_void_p, c_void_p],
    "cm_nn_weights": [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p],
    "cm_nn_adam": [c_void_p, c_ll, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_float, c_float, c_int],
    "cm_nn_colsum": [c_void_p, c_ll, c_int, c_void_p, c_void_p],
    "cm_nn_relu_bwd": [c_void_p, c_ll, c_void_p, c_void_p, c_void_p, c_void_p],
    "cm_nn_pool": [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p],
    "cm_nn_pool_bwd": [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p],
    "cm_nn_up": [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p],
    "cm_nn_up_bwd": [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p],
    "cm_nn_mse": [c_void_p, c_ll, c_int, c_int, c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_void_p],
    "cm_nn_features": [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_void_p,
                       c_void_p, c_void_p, c_float, c_void_p],
    "cm_nn_node_to_cell": [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_void_p],
    "cm_nn_cell_to_node": [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_ll,
                           c_void_p, c_double, c_void_p],
    "cm_nn_node_groups": [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_ll, c_void_p],
    "cm_nn_group_centres": [c_void_p, c_int, c_int, c_int, c_ll, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                            c_int, c_void_p],
    "cm_nn_zt": [c_void_p, c_int, c_int, c_int, c_ll, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                 c_void_p, c_void_p, c_void_p, c_void_p],
    "cm_nn_zc": [c_void_p, c_int, c_int, c_int, c_ll, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                 c_void_p, c_void_p, c_void_p, c_void_p],
    "cm_nn_fill": [c_void_p, c_ll, c_void_p, c_float],
    "cm_nn_cast": [c_void_p, c_ll, c_void_p, c_void_p, c_int],
    "cm_nn_axpy": [c_void_p, c_ll, c_float, c_void_p, c_void_p],
}
_RET = {"cm_last_error": ctypes.c_char_p, "cm_version": c_int, "cm_nn_last_error": ctypes.c_char_p}

_lib = None


def load(path=LIB_PATH):
    """Load the library (once).  Raises OSError with a clear message when it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with `python -m comech_amd.build` (hipcc, gfx950); "
                      "the solver has no CPU fallback")
    lib = ctypes.CDLL(path)
    for name, args in list(SIGNATURES.items()) + list(SIGNATURES_NN.items()):
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = _RET.get(name, c_int)
    _lib = lib
    return lib


class CmError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise CmError(load().cm_last_error().decode())
    return rc


def check_nn(rc):
    """Status of a cm_nn_* call (include/comech_learned.h)."""
    if rc != 0:
        raise CmError(load().cm_nn_last_error().decode())
    return rc


def ptr(a, ctype):
    """Host numpy array -> ctypes pointer (array must stay alive during the call)."""
    return a.ctypes.data_as(ctypes.POINTER(ctype))
