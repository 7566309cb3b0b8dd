"""ctypes binding of libdachshund_gpu.so (include/dachshund_gpu.h).

The library is built in-tree by ``make -C dachshund_amd``.  Loading fails
loudly when it is missing: there is no host fallback for any GPU routine.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdachshund_gpu.so")

DS_OK, DS_E_ARG, DS_E_HIP, DS_E_NOMEM, DS_E_STATE = range(5)
DS_FORM_SN, DS_FORM_W = 0, 1
DS_PREC_F64, DS_PREC_F32 = 0, 1
DS_F32_VALU, DS_F32_MFMA, DS_F32_MFMA_CLAMPED = 0, 1, 2
DS_PRECOND_JACOBI, DS_PRECOND_BLOCK = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p


class DsCovar(C.Structure):
    _fields_ = [("var_s", C.c_double), ("c_perp", C.c_double),
                ("c_para", C.c_double), ("dx_inv", C.c_double),
                ("table_n", C.c_int), ("table", _dp)]


class DsPcgParams(C.Structure):
    _fields_ = [("n", C.c_int), ("max_iter", C.c_int), ("step_r", C.c_int),
                ("tol", C.c_double), ("verbose", C.c_int)]


class DsPcgResult(C.Structure):
    _fields_ = [("success", C.c_int), ("num_iters", C.c_int),
                ("n_norms", C.c_int), ("n_true", C.c_int),
                ("n_matvec64", C.c_int), ("n_matvec32", C.c_int),
                ("norm_b", C.c_double), ("final_norm", C.c_double),
                ("true_norm", C.c_double), ("seconds", C.c_double)]


class DsProblemStats(C.Structure):
    _fields_ = [("n", C.c_int), ("n_blocks", C.c_int), ("n_tiles", C.c_int),
                ("tile_pairs", C.c_longlong),
                ("pairs_per_matvec", C.c_longlong),
                ("sub_pairs_per_matvec", C.c_longlong),
                ("setup_seconds", C.c_double),
                ("sub_pairs_evaluated", C.c_longlong),
                ("f64_pairs_per_matvec", C.c_longlong)]


# name -> (restype, argtypes); every symbol of include/dachshund_gpu.h
SIGNATURES = {
    "ds_plan_tiles": (C.c_int, [C.c_int, _dp, C.c_int, C.POINTER(DsCovar),
                                _ip, _ip, _ip, C.c_int, _ip, _ip]),
    "ds_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "ds_ctx_destroy": (C.c_int, [_vp]),
    "ds_ctx_last_error": (C.c_char_p, [_vp]),
    "ds_ctx_synchronize": (C.c_int, [_vp]),
    "ds_ctx_num_cu": (C.c_int, [_vp]),
    "ds_ctx_set_host_geometry": (C.c_int, [_vp, C.c_int]),
    "ds_problem_create": (C.c_int, [_vp, C.c_int, C.c_int, _dp,
                                    C.POINTER(DsCovar), C.POINTER(_vp)]),
    "ds_problem_destroy": (C.c_int, [_vp]),
    "ds_problem_get_stats": (C.c_int, [_vp, C.POINTER(DsProblemStats)]),
    "ds_problem_get_geometry": (C.c_int, [_vp, C.c_int, _vp, C.c_longlong,
                                          C.POINTER(C.c_longlong)]),
    "ds_problem_set_waves": (C.c_int, [_vp, C.c_int]),
    "ds_batch_create": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _dp,
                                  C.POINTER(DsCovar), C.POINTER(_vp)]),
    "ds_batch_num_sub": (C.c_int, [_vp]),
    "ds_batch_sub_pairs": (C.c_int, [_vp, C.POINTER(C.c_longlong),
                                     C.POINTER(C.c_longlong)]),
    "ds_batch_solve": (C.c_int, [_vp, C.POINTER(DsPcgParams), C.c_int,
                                 C.c_double, C.c_int,
                                 C.POINTER(DsPcgResult)]),
    "ds_problem_set_launch": (C.c_int, [_vp, C.c_int, C.c_int]),
    "ds_problem_set_f32_kernel": (C.c_int, [_vp, C.c_int]),
    "ds_problem_set_precond": (C.c_int, [_vp, C.c_int]),
    "ds_problem_set_fused": (C.c_int, [_vp, C.c_int]),
    "ds_problem_build_dense_bf16": (C.c_int, [_vp]),
    "ds_problem_set_dense": (C.c_int, [_vp, C.c_int]),
    "ds_problem_dense_product": (C.c_int, [_vp, _dp, _dp]),
    "ds_problem_time_dense": (C.c_int, [_vp, C.c_int, _dp]),
    "ds_problem_set_tile": (C.c_int, [_vp, C.c_int]),
    "ds_problem_get_launch": (C.c_int, [_vp, _ip, _ip, _ip, _ip]),
    "ds_problem_autotune": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _ip,
                                      _dp]),
    "ds_A_product": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "ds_A_residual": (C.c_int, [_vp, _dp, _dp]),
    "ds_precond": (C.c_int, [_vp, _dp, _dp]),
    "ds_precond_r32": (C.c_int, [_vp, _dp, _dp]),
    "ds_norm_b": (C.c_int, [_vp, _dp]),
    "ds_solve_pcg": (C.c_int, [_vp, C.POINTER(DsPcgParams), _dp, _dp,
                               C.c_int, C.POINTER(DsPcgResult)]),
    "ds_solve_mixed": (C.c_int, [_vp, C.POINTER(DsPcgParams), C.c_double,
                                 _dp, _dp, C.c_int,
                                 C.POINTER(DsPcgResult)]),
    "ds_problem_set_x": (C.c_int, [_vp, _dp]),
    "ds_problem_get_x": (C.c_int, [_vp, _dp]),
    "ds_solve_pcg_resident": (C.c_int, [_vp, C.POINTER(DsPcgParams),
                                        C.POINTER(DsPcgResult)]),
    "ds_solve_mixed_resident": (C.c_int, [_vp, C.POINTER(DsPcgParams),
                                          C.c_double,
                                          C.POINTER(DsPcgResult)]),
    "ds_problem_get_norms": (C.c_int, [_vp, _dp, C.c_int, _ip]),
    "ds_problem_get_deltas": (C.c_int, [_vp, _dp, C.c_int, _ip]),
    "ds_problem_fill_x": (C.c_int, [_vp, C.c_double]),
    "ds_problem_time_matvec": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "ds_solve_cf": (C.c_int, [_vp, C.c_int, _dp]),
    "ds_map_covar_diag": (C.c_int, [_vp, C.c_int, _dp, C.c_int, _dp,
                                    C.POINTER(DsCovar),
                                    C.POINTER(DsPcgParams), _dp, _ip, _ip]),
    "ds_ws_num_params": (C.c_int, []),
    "ds_ws_feature_dim": (C.c_int, []),
    "ds_ws_create": (C.c_int, [_vp, C.c_ulonglong, C.POINTER(_vp)]),
    "ds_ws_destroy": (C.c_int, [_vp]),
    "ds_ws_set_params": (C.c_int, [_vp, _fp, C.c_int]),
    "ds_ws_get_params": (C.c_int, [_vp, _fp, C.c_int]),
    "ds_ws_set_quantized": (C.c_int, [_vp, C.c_int]),
    "ds_ws_features": (C.c_int, [_vp, _vp, _fp]),
    "ds_ws_predict": (C.c_int, [_vp, _vp]),
    "ds_ws_add_sample": (C.c_int, [_vp, _vp, C.c_ulonglong]),
    "ds_ws_num_samples": (C.c_int, [_vp, C.POINTER(C.c_longlong)]),
    "ds_ws_clear_samples": (C.c_int, [_vp]),
    "ds_ws_train": (C.c_int, [_vp, C.c_int, C.c_int, C.c_float, C.c_float,
                              C.c_ulonglong, _dp]),
    "ds_ws_eval_batch": (C.c_int, [_vp, C.c_int, _fp, _fp, _fp, _fp, _dp]),
    "ds_problem_smp_grid": (C.c_int, [_vp, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_int, C.c_int,
                                      _dp]),
    "ds_problem_smp_grid_noalloc": (C.c_int, [_vp, C.c_double, C.c_double,
                                              C.c_double, C.c_int, C.c_int,
                                              C.c_int, _dp]),
    "ds_smp_product": (C.c_int, [_vp, C.c_int, _dp, C.c_int, _dp,
                                 C.POINTER(DsCovar), _dp, _dp]),
    "ds_smp_product_grid": (C.c_int, [_vp, C.c_int, _dp, C.c_double,
                                      C.c_double, C.c_double, C.c_int,
                                      C.c_int, C.c_int, C.POINTER(DsCovar),
                                      _dp, _dp]),
    "ds_sigma_m_noise_dom": (C.c_int, [_vp, C.c_int, _dp, C.c_int, _dp,
                                       C.POINTER(DsCovar), _dp]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libdachshund_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with "
                "`make -C dachshund_amd` (or __graft_entry__.build()); the "
                "GPU path has no fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class DsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ds error {code}: {msg}")
        self.code = code


def check(rc, ctx=None, what=""):
    if rc != DS_OK:
        msg = what
        if ctx is not None:
            m = lib().ds_ctx_last_error(ctx)
            if m:
                msg = f"{what}: {m.decode()}" if what else m.decode()
        raise DsError(rc, msg)
