"""ctypes binding of libsp_hip.so (the C ABI declared in include/stencilparareal.h).

The product path has no fallback: if the library is not built, or a call fails,
an exception is raised.  Nothing here touches the CPU oracle.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsp_hip.so")

SP_OK = 0
SP_ERR_INVALID = 1
SP_ERR_HIP = 2
SP_ERR_UNSUPPORTED = 3
SP_ERR_DIVERGED = 4
SP_ERR_NOMEM = 5

SP_F64, SP_F32, SP_BF16 = 0, 1, 2
SP_COARSE_EULER, SP_COARSE_RK4, SP_COARSE_RK4_SPECTRAL, SP_COARSE_EULER_SPECTRAL = 0, 1, 2, 3
SP_PATH_FUSED, SP_PATH_STAGED = 0, 1
SP_KERNEL_AUTO, SP_KERNEL_SCALAR, SP_KERNEL_PAIR, SP_KERNEL_QUAD, SP_KERNEL_SPLIT = -1, 0, 1, 2, 3
SP_MAX_ITERS = 64

DTYPE_CODES = {"f64": SP_F64, "f32": SP_F32, "bf16": SP_BF16}
DTYPE_SIZES = {SP_F64: 8, SP_F32: 4, SP_BF16: 2}

# Every symbol the header declares (checked by the ABI tests).
EXPORTS = [
    "sp_version", "sp_last_error", "sp_ctx_create", "sp_ctx_destroy", "sp_ctx_set_stream",
    "sp_ctx_get_stream", "sp_ctx_sync", "sp_malloc", "sp_free", "sp_memcpy_h2d", "sp_memcpy_d2h",
    "sp_layout_compact", "sp_layout_stella_cuda", "sp_layout_storage_elems", "sp_prop_create",
    "sp_prop_destroy", "sp_prop_set_upwind", "sp_prop_set_path", "sp_prop_set_tile", "sp_prop_get_tile", "sp_prop_autotune", "sp_do_rk4",
    "sp_do_euler", "sp_do_rk4_timestep", "sp_do_euler_timestep", "sp_do_rk4_ex", "sp_do_rk4_batched",
    "sp_do_rk4_noalloc",
    "sp_periodic_fill",
    "sp_parareal_update", "sp_norms", "sp_fill_exact", "sp_fill_gaussians", "sp_parareal_create",
    "sp_parareal_destroy", "sp_parareal_begin", "sp_parareal_iterate", "sp_parareal_iteration_times",
    "sp_parareal_solution",
    "sp_parareal_run", "sp_parareal_serial",
    "sp_batch_create", "sp_batch_destroy", "sp_batch_run", "sp_batch_increments", "sp_batch_iteration_times",
    "sp_batch_problem_info",
    "sp_batch_serial",
]


class SPError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[sp error {code}] {msg}")
        self.code = code


class Layout(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int), ("halo", C.c_int),
                ("sj", C.c_longlong), ("sk", C.c_longlong), ("origin", C.c_longlong),
                ("dtype", C.c_int)]


class PararealCfg(C.Structure):
    _fields_ = [("slices", C.c_int), ("kmax", C.c_int), ("end_time", C.c_double),
                ("steps_fine", C.c_int), ("steps_coarse", C.c_int), ("coarse_kind", C.c_int),
                ("coarse_dtype", C.c_int), ("coarse_substeps", C.c_int), ("tol", C.c_double),
                ("async_schedule", C.c_int), ("use_graphs", C.c_int), ("autotune", C.c_int),
                ("delta_correction", C.c_int), ("initial_fp64", C.c_int)]


class PararealReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("steps_fine_per_slice", C.c_int),
                ("steps_coarse_per_slice", C.c_int), ("dt_fine", C.c_double),
                ("dt_coarse", C.c_double), ("increment", C.c_double * SP_MAX_ITERS),
                ("t_init", C.c_double), ("t_fine", C.c_double * SP_MAX_ITERS),
                ("t_coarse", C.c_double * SP_MAX_ITERS), ("t_total", C.c_double),
                ("ms_fine_events", C.c_double * SP_MAX_ITERS), ("t_update", C.c_double * SP_MAX_ITERS)]


class Problem(C.Structure):
    _fields_ = [("n", C.c_int), ("slices", C.c_int), ("nu", C.c_double), ("cx", C.c_double),
                ("cy", C.c_double), ("cz", C.c_double), ("end_time", C.c_double),
                ("steps_fine", C.c_int), ("steps_coarse", C.c_int)]


class BatchCfg(C.Structure):
    _fields_ = [("kmax", C.c_int), ("coarse_dtype", C.c_int), ("coarse_substeps", C.c_int),
                ("tol", C.c_double), ("use_graphs", C.c_int), ("autotune", C.c_int)]


class BatchReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("increment", C.c_double * SP_MAX_ITERS), ("t_init", C.c_double),
                ("t_total", C.c_double), ("lattice_updates", C.c_double)]


class BatchInfo(C.Structure):
    _fields_ = [("steps_fine_per_slice", C.c_int), ("steps_coarse_per_slice", C.c_int),
                ("dt_fine", C.c_double), ("dt_coarse", C.c_double), ("fine_variant", C.c_int),
                ("fine_tx", C.c_int), ("fine_ty", C.c_int), ("fine_nthreads", C.c_int), ("fine_kchunk", C.c_int)]


_vp = C.c_void_p
_lp = C.POINTER(Layout)
_dp = C.POINTER(C.c_double)

_SIGS = {
    "sp_version": (C.c_char_p, []),
    "sp_last_error": (C.c_char_p, []),
    "sp_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "sp_ctx_destroy": (C.c_int, [_vp]),
    "sp_ctx_set_stream": (C.c_int, [_vp, _vp]),
    "sp_ctx_get_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
    "sp_ctx_sync": (C.c_int, [_vp]),
    "sp_malloc": (C.c_int, [_vp, C.c_longlong, C.POINTER(_vp)]),
    "sp_free": (C.c_int, [_vp, _vp]),
    "sp_memcpy_h2d": (C.c_int, [_vp, _vp, _vp, C.c_longlong]),
    "sp_memcpy_d2h": (C.c_int, [_vp, _vp, _vp, C.c_longlong]),
    "sp_layout_compact": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _lp]),
    "sp_layout_stella_cuda": (C.c_int, [C.c_int, C.c_int, C.c_int, _lp]),
    "sp_layout_storage_elems": (C.c_longlong, [_lp]),
    "sp_prop_create": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int] + [C.c_double] * 6 + [C.POINTER(_vp)]),
    "sp_prop_destroy": (C.c_int, [_vp]),
    "sp_prop_set_upwind": (C.c_int, [_vp, C.c_int]),
    "sp_prop_set_path": (C.c_int, [_vp, C.c_int]),
    "sp_prop_set_tile": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "sp_prop_get_tile": (C.c_int, [_vp, C.c_int, C.c_int] + [C.POINTER(C.c_int)] * 5),
    "sp_prop_autotune": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp]),
    "sp_do_rk4": (C.c_int, [_vp, _lp, _vp, _lp, _vp, C.c_double, C.c_double, C.c_int]),
    "sp_do_rk4_noalloc": (C.c_int, [_vp, _lp, _vp, _lp, _vp, C.c_double, C.c_double, C.c_int]),
    "sp_do_euler": (C.c_int, [_vp, _lp, _vp, _lp, _vp, C.c_double, C.c_double, C.c_int]),
    "sp_do_rk4_timestep": (C.c_int, [_vp, _lp, _vp, _lp, _vp, C.c_double, C.c_double]),
    "sp_do_euler_timestep": (C.c_int, [_vp, _lp, _vp, _lp, _vp, C.c_double, C.c_double]),
    "sp_do_rk4_ex": (C.c_int, [_vp, _lp, _vp, _lp, _vp, C.c_double, C.c_double, C.c_int]),
    "sp_do_rk4_batched": (C.c_int, [_vp, C.c_int, _lp, _vp, _lp, _vp, C.c_longlong, C.c_double, C.c_double,
                                    C.c_int]),
    "sp_periodic_fill": (C.c_int, [_vp, _lp, _vp]),
    "sp_parareal_update": (C.c_int, [_vp, _lp, _vp, _vp, _vp, _vp]),
    "sp_norms": (C.c_int, [_vp, _lp, _vp, _lp, _vp, _dp]),
    "sp_fill_exact": (C.c_int, [_vp, _lp, _vp] + [C.c_double] * 6),
    "sp_fill_gaussians": (C.c_int, [_vp, _lp, _vp, C.c_int, _dp, _dp, _dp, _dp, _dp]),
    "sp_parareal_create": (C.c_int, [_vp, C.POINTER(PararealCfg), C.POINTER(_vp)]),
    "sp_parareal_destroy": (C.c_int, [_vp]),
    "sp_parareal_begin": (C.c_int, [_vp, _lp, _vp]),
    "sp_parareal_iterate": (C.c_int, [_vp, _dp]),
    "sp_parareal_iteration_times": (C.c_int, [_vp, _dp, _dp, _dp]),
    "sp_parareal_solution": (C.c_int, [_vp, _lp, _vp]),
    "sp_parareal_run": (C.c_int, [_vp, _lp, _vp, _lp, _vp, C.POINTER(PararealReport)]),
    "sp_parareal_serial": (C.c_int, [_vp, _lp, _vp, _lp, _vp]),
    "sp_batch_create": (C.c_int, [_vp, C.c_int, C.POINTER(Problem), C.POINTER(BatchCfg), C.POINTER(_vp)]),
    "sp_batch_destroy": (C.c_int, [_vp]),
    "sp_batch_run": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(BatchReport)]),
    "sp_batch_increments": (C.c_int, [_vp, C.c_int, _dp]),
    "sp_batch_iteration_times": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "sp_batch_problem_info": (C.c_int, [_vp, C.c_int, C.POINTER(BatchInfo)]),
    "sp_batch_serial": (C.c_int, [_vp, C.c_int, _vp, _vp]),
}

_lib = None


def load(path: str | None = None):
    """Loads libsp_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"{p} is missing: build it with `python -m stencilparareal_amd.build` "
            "(the HIP library is required; there is no CPU fallback)")
    L = C.CDLL(p)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = L
    return L


def check(rc: int):
    if rc != SP_OK:
        msg = load().sp_last_error()
        raise SPError(rc, msg.decode() if msg else "")
    return rc


def call(name: str, *args):
    return check(getattr(load(), name)(*args))
