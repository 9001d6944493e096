"""ctypes binding of libpawpy_amd.so (the HIP library).

The library is built in-tree by ``python -m pawpyseed_amd.build``; loading
fails loudly when it is missing -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_HERE, "libpawpy_amd.so")

PP_OK, PP_ERR_ARG, PP_ERR_HIP, PP_ERR_NOMEM, PP_ERR_UNSUPPORTED, PP_ERR_IO = range(6)
PP_FP64, PP_FP32, PP_BF16X3, PP_BF16X6, PP_FP16X3 = 0, 1, 2, 3, 4
PRECISIONS = {"fp64": PP_FP64, "fp32": PP_FP32, "bf16x3": PP_BF16X3, "bf16x6": PP_BF16X6, "fp16x3": PP_FP16X3}


class PAWpyError(Exception):
    """Error raised by the projection path (mirrors pawpyseed.core.utils.PAWpyError)."""


class LibraryMissing(PAWpyError, ImportError):
    pass


c_int_p = ct.POINTER(ct.c_int)
c_double_p = ct.POINTER(ct.c_double)
c_void_p = ct.c_void_p

_lib = None

# name -> (restype, argtypes)
SIGNATURES = {
    "pp_last_error": (ct.c_char_p, []),
    "pp_last_status": (ct.c_int, []),
    "pp_version": (ct.c_char_p, []),
    "pp_set_precision": (ct.c_int, [ct.c_int]),
    "pp_get_precision": (ct.c_int, []),
    "read_wavefunctions": (c_void_p, [ct.c_char_p, c_double_p]),
    "read_wavefunctions_from_str": (c_void_p, [ct.c_char_p, c_double_p]),
    "pp_pswf_from_arrays": (c_void_p, [ct.c_int, ct.c_int, ct.c_int, ct.c_double, c_double_p, c_double_p,
                                       c_double_p, c_int_p, c_int_p, c_void_p, c_double_p, c_double_p]),
    "free_pswf": (None, [c_void_p]),
    "free_ppot_list": (None, [c_void_p, ct.c_int]),
    "get_occs": (c_double_p, [c_void_p]),
    "get_nband": (ct.c_int, [c_void_p]),
    "get_nwk": (ct.c_int, [c_void_p]),
    "get_nspin": (ct.c_int, [c_void_p]),
    "is_ncl": (ct.c_int, [c_void_p]),
    "get_encut": (ct.c_double, [c_void_p]),
    "get_energy": (ct.c_double, [c_void_p, ct.c_int, ct.c_int, ct.c_int]),
    "get_occ": (ct.c_double, [c_void_p, ct.c_int, ct.c_int, ct.c_int]),
    "set_num_sites": (None, [c_void_p, ct.c_int]),
    "free_ptr": (None, [c_void_p]),
    "get_projector_list": (c_void_p, [ct.c_int, c_int_p, c_int_p, c_double_p, c_double_p, c_double_p,
                                      c_double_p, c_double_p, ct.c_double]),
    "setup_projections": (ct.c_int, [c_void_p, c_void_p, ct.c_int, ct.c_int, c_int_p, c_int_p, c_double_p]),
    "pseudoprojection": (ct.c_int, [c_void_p, c_void_p, c_void_p, ct.c_int, ct.c_int]),
    "overlap_setup_real": (ct.c_int, [c_void_p, c_void_p, c_int_p, c_int_p, c_double_p, c_double_p, c_int_p, c_int_p,
                                      c_int_p, c_int_p, ct.c_int, ct.c_int, ct.c_int]),
    "compensation_terms": (ct.c_int, [c_void_p, ct.c_int, c_void_p, c_void_p, ct.c_int, ct.c_int, ct.c_int,
                                      ct.c_int, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p,
                                      c_double_p, c_int_p, c_double_p, c_int_p, ct.c_int]),
    "overlap_setup_recip": (ct.c_int, [c_void_p, c_void_p, c_int_p, c_int_p, c_double_p, c_double_p, c_int_p,
                                       c_int_p, c_int_p, c_int_p, ct.c_int, ct.c_int, ct.c_int]),
    "compensation_terms_recip": (ct.c_int, [c_void_p, ct.c_int, c_void_p, c_void_p, ct.c_int, ct.c_int, ct.c_int,
                                            ct.c_int, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p,
                                            c_double_p, c_int_p, c_double_p, c_int_p, ct.c_int]),
    "pp_aug_freqs_batched": (ct.c_int, [c_void_p, c_void_p, ct.c_int, c_double_p, c_int_p, c_void_p, c_void_p,
                                        ct.c_long, c_int_p, ct.c_int, c_void_p, ct.c_long, ct.c_int]),
    "fft3d": (ct.c_int, [c_void_p, c_int_p, c_double_p, c_double_p, c_int_p, c_void_p, ct.c_int, c_int_p]),
    "pp_get_projections": (ct.c_int, [c_void_p, ct.c_int, ct.c_int, c_void_p]),
    "realspace_state": (ct.c_int, [c_void_p, ct.c_int, ct.c_int, c_void_p, c_int_p, c_int_p, c_double_p]),
    "ncl_realspace_state": (ct.c_int, [c_void_p, ct.c_int, ct.c_int, c_void_p, c_int_p, c_int_p, c_double_p]),
    "remove_phase": (ct.c_int, [c_void_p, ct.c_int, c_void_p, c_int_p]),
    "ae_state_density": (ct.c_int, [c_double_p, ct.c_int, ct.c_int, c_void_p, c_int_p, c_int_p, c_double_p]),
    "ae_chg_density": (ct.c_int, [c_double_p, c_void_p, c_int_p, c_int_p, c_double_p]),
    "ncl_ae_chg_density": (ct.c_int, [c_double_p, c_void_p, c_int_p, c_int_p, c_double_p]),
    "project_realspace_state": (ct.c_int, [c_void_p, ct.c_int, c_void_p, c_void_p, c_int_p, c_int_p, c_double_p,
                                           c_int_p, c_double_p]),
    "pp_rs_plan_create": (ct.c_int, [c_void_p, ct.POINTER(c_void_p), c_double_p, c_int_p, ct.c_int, c_int_p,
                                     c_double_p, ct.c_int, c_int_p, c_int_p, c_int_p, c_double_p, c_double_p]),
    "pp_rs_plan_destroy": (None, [c_void_p]),
    "pp_rs_plan_info": (ct.c_int, [c_void_p, c_int_p, ct.POINTER(ct.c_long), c_int_p, c_int_p, c_int_p]),
    "pp_realspace_batched": (ct.c_int, [c_void_p, c_void_p, c_double_p, ct.c_int, c_void_p, c_void_p, ct.c_long,
                                        ct.c_int, c_void_p, ct.c_long, ct.c_int, c_void_p, ct.c_long, c_double_p,
                                        c_void_p, ct.c_int]),
    "pp_ctx_create": (ct.c_int, [ct.POINTER(c_void_p), ct.c_int]),
    "pp_ctx_destroy": (None, [c_void_p]),
    "pp_ctx_set_stream": (ct.c_int, [c_void_p, c_void_p]),
    "pp_ctx_synchronize": (ct.c_int, [c_void_p]),
    "pp_ctx_set_timing": (ct.c_int, [c_void_p, ct.c_int]),
    "pp_ctx_timing_read": (ct.c_int, [c_void_p, ct.c_int, ct.POINTER(ct.c_char_p), c_double_p,
                                      ct.POINTER(ct.c_long)]),
    "pp_ctx_timing_reset": (ct.c_int, [c_void_p]),
    "pp_ctx_set_tuning": (ct.c_int, [c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int]),
    "pp_ctx_set_param": (ct.c_int, [c_void_p, ct.c_char_p, ct.c_int]),
    "pp_ctx_get_param": (ct.c_int, [c_void_p, ct.c_char_p, c_int_p]),
    "pp_graph_capture": (ct.c_int, [c_void_p, c_void_p, ct.c_int, c_double_p, c_int_p, c_void_p, c_void_p,
                                    ct.c_long, ct.c_int, c_void_p, ct.c_int, ct.POINTER(c_void_p)]),
    "pp_graph_replay": (ct.c_int, [c_void_p]),
    "pp_graph_destroy": (None, [c_void_p]),
    "pp_autotune": (ct.c_int, [c_void_p, c_void_p, ct.c_int, c_double_p, c_int_p, c_void_p, c_void_p, ct.c_long,
                               ct.c_int, c_void_p, ct.c_int, ct.c_int, ct.c_char_p, ct.c_int]),
    "pp_sphere_create": (ct.c_int, [c_void_p, ct.POINTER(c_void_p), c_double_p, c_int_p, ct.c_int, c_int_p,
                                    c_double_p, ct.c_int, c_int_p, c_int_p, c_double_p, c_int_p, c_double_p,
                                    ct.c_int]),
    "pp_sphere_destroy": (None, [c_void_p]),
    "pp_sphere_info": (ct.c_int, [c_void_p, c_int_p, ct.POINTER(ct.c_long), c_int_p]),
    "pp_sphere_coverage": (ct.c_int, [c_void_p, ct.POINTER(ct.c_long)]),
    "pp_sphere_fused_info": (ct.c_int, [c_void_p, c_int_p, c_int_p, c_int_p, ct.POINTER(ct.c_long), c_int_p]),
    "pp_sphere_rebuild": (ct.c_int, [c_void_p, c_double_p]),
    "pp_setup_projections_batched": (ct.c_int, [c_void_p, c_void_p, ct.c_int, c_double_p, c_int_p, c_void_p,
                                                c_void_p, ct.c_long, ct.c_int, c_void_p, ct.c_int]),
    "pp_overlap_gemm": (ct.c_int, [c_void_p, ct.c_int, ct.c_int, ct.c_int, c_int_p, c_void_p, ct.c_long, ct.c_int,
                                   c_void_p, ct.c_long, ct.c_int, c_void_p, ct.c_int]),
    "pp_setup_projections_multi": (ct.c_int, [c_void_p, ct.c_int, c_void_p, ct.c_int, c_double_p, c_int_p,
                                              c_void_p, c_int_p, c_void_p, ct.POINTER(ct.c_long), c_void_p,
                                              ct.c_int]),
    "pp_graph_capture_multi": (ct.c_int, [c_void_p, ct.c_int, c_void_p, ct.c_int, c_double_p, c_int_p, c_void_p,
                                          c_int_p, c_void_p, ct.POINTER(ct.c_long), c_void_p, ct.c_int,
                                          ct.POINTER(c_void_p)]),
    "pp_autotune_multi": (ct.c_int, [c_void_p, ct.c_int, c_void_p, ct.c_int, c_double_p, c_int_p, c_void_p,
                                     c_int_p, c_void_p, ct.POINTER(ct.c_long), c_void_p, ct.c_int, ct.c_int,
                                     ct.c_char_p, ct.c_int]),
    "pp_overlap_gemm_noalloc": (ct.c_int, [c_void_p, ct.c_int, ct.c_int, ct.c_int, c_int_p, c_void_p, ct.c_long,
                                           ct.c_int, c_void_p, ct.c_long, ct.c_int, c_void_p, ct.c_int]),
    "pp_noalloc_scratch_bytes": (ct.c_size_t, []),
    "pp_augmentation": (ct.c_int, [c_void_p, ct.c_int, ct.c_int, ct.c_int, c_void_p, ct.c_int, ct.c_int, c_int_p,
                                   c_int_p, c_void_p, ct.c_int, ct.c_int, c_int_p, c_int_p, ct.c_int, c_int_p,
                                   c_int_p, ct.c_int, c_int_p, c_int_p, c_double_p, c_void_p]),
    "pp_fft3d_batched": (ct.c_int, [c_void_p, c_double_p, c_int_p, ct.c_int, c_void_p, c_void_p, ct.c_long,
                                    ct.c_int, c_void_p, ct.c_int]),
    "pp_augmentation_offsite": (ct.c_int, [c_void_p, ct.c_int, ct.c_int, ct.c_int, c_void_p, ct.c_int, ct.c_int,
                                           c_int_p, c_void_p, ct.c_int, ct.c_int, c_int_p, ct.c_int, c_int_p,
                                           c_int_p, c_void_p, ct.POINTER(ct.c_long), c_double_p, c_double_p,
                                           c_void_p]),
    "pp_sbt_forward": (ct.c_int, [ct.c_double, ct.c_double, ct.c_int, ct.c_int, c_double_p, c_double_p, ct.c_int,
                                  c_double_p, c_double_p]),
    "pp_smooth_diffwave": (ct.c_int, [ct.c_int, c_double_p, c_double_p, ct.c_int, ct.c_int, ct.c_double, ct.c_int,
                                      c_double_p, c_double_p, c_double_p]),
    "pp_offsite_overlap": (ct.c_int, [c_double_p, ct.c_int, c_double_p, c_double_p, ct.c_int, c_double_p,
                                      c_double_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, c_double_p]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(f"{LIB_PATH} is not built; run `python -m pawpyseed_amd.build`")
        L = ct.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int, what: str = ""):
    if status != PP_OK:
        msg = lib().pp_last_error().decode(errors="replace")
        raise PAWpyError(f"{what or 'pawpy_amd'} failed (status {status}): {msg}")
    return status


def ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def iptr(a):
    return None if a is None else np.ascontiguousarray(a, np.int32).ctypes.data_as(c_int_p)


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
