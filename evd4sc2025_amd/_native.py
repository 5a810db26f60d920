"""ctypes binding of libevd4sc.so (the C ABI declared in include/evd4sc.h).

The library is loaded from this package directory only; a missing or stale
build raises immediately - there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libevd4sc.so"

_lib = None

_vp, _l, _i, _f, _d, _ch, _u64, _sz = (C.c_void_p, C.c_long, C.c_int, C.c_float, C.c_double, C.c_char,
                                       C.c_ulonglong, C.c_size_t)

# name -> (restype, argtypes)
SIGNATURES = {
    "evd_ctx_create": (_i, [_i, _vp, C.POINTER(_vp)]),
    "evd_ctx_destroy": (_i, [_vp]),
    "evd_ctx_set_stream": (_i, [_vp, _vp]),
    "evd_ctx_get_stream": (_vp, [_vp]),
    "evd_ctx_last_error": (C.c_char_p, [_vp]),
    "evd_ctx_reserve": (_i, [_vp, _sz]),
    "evd_ctx_arena_bytes": (_sz, [_vp]),
    "evd_ctx_set_budget": (_i, [_vp, _sz]),
    "evd_ctx_set_tuning": (_i, [_vp, _i, _i, _i, _i]),
    "evd_ctx_set_knob": (_i, [_vp, C.c_char_p, _i]),
    "evd_ctx_set_precision": (_i, [_vp, _i]),
    "evd_ctx_enable_timing": (_i, [_vp, _i]),
    "evd_ctx_stage_ms": (_i, [_vp, C.POINTER(_f), _i]),
    "evd_ctx_synchronize": (_i, [_vp]),
    "evd_panel_qr_f32": (_i, [_vp, _l, _l, _vp, _l, _vp, _l, _vp, _l]),
    "evd_orgqr_f32": (_i, [_vp, _l, _l, _vp, _l, _vp, _l, _vp]),
    "evd_syr2k_f32": (_i, [_vp, _l, _l, _f, _vp, _l, _vp, _l, _f, _vp, _l]),
    "evd_sy2sb_f32": (_i, [_vp, _l, _l, _l, _vp, _l, _vp, _l, _vp, _l]),
    "evd_band_pack_f32": (_i, [_vp, _l, _l, _vp, _l, _vp, _l]),
    "evd_sb2st_f64": (_i, [_vp, _l, _l, _vp, _l, _vp, _l, _vp, _vp]),
    "evd_sb2st_f64_noalloc": (_i, [_vp, _l, _l, _vp, _l, _vp, _l, _vp, _vp]),
    "evd_stedc_f64": (_i, [_vp, _l, _vp, _vp, _vp, _l]),
    "evd_wy_from_slide_f32": (_i, [_vp, _l, _l, _l, _vp, _l, _vp, _l]),
    "evd_q_from_wy_f32": (_i, [_vp, _l, _l, _vp, _l, _vp, _l, _vp, _l]),
    "evd_bc_back_f32": (_i, [_vp, _l, _l, _l, _vp, _l, _vp, _l, _i]),
    "evd_sbr_back_f32": (_i, [_vp, _l, _l, _l, _vp, _l, _vp, _l, _vp, _l]),
    "evd_syevd_f32": (_i, [_vp, _ch, _l, _l, _l, _vp, _l, _vp, _vp, _l]),
    "evd_panel_qr_f64": (_i, [_vp, _l, _l, _vp, _l, _vp, _l, _vp, _l]),
    "evd_orgqr_f64": (_i, [_vp, _l, _l, _vp, _l, _vp, _l, _vp]),
    "evd_syr2k_f64": (_i, [_vp, _l, _l, _d, _vp, _l, _vp, _l, _d, _vp, _l]),
    "evd_sy2sb_f64": (_i, [_vp, _l, _l, _l, _vp, _l, _vp, _l, _vp, _l]),
    "evd_wy_from_slide_f64": (_i, [_vp, _l, _l, _l, _vp, _l, _vp, _l]),
    "evd_q_from_wy_f64": (_i, [_vp, _l, _l, _vp, _l, _vp, _l, _vp, _l]),
    "evd_bc_back_f64": (_i, [_vp, _l, _l, _l, _vp, _l, _vp, _l, _i]),
    "evd_syevd_workspace_f32": (_sz, [_l, _l, _l, _ch]),
    "evd_syevd_workspace_ex": (_sz, [_l, _l, _l, _ch, _i]),
    "evd_autotune": (_i, [_vp, _l, _l, _i, C.POINTER(_i)]),
    "evd_syevd_batched_f32": (_i, [_vp, _ch, _i, _l, _l, _l, _vp, _l, _l, _vp, _l, _vp, _l, _l]),
    "evd_syevd_batched_workspace_f32": (_sz, [_i, _l, _l, _l, _ch, _i]),
    "evd_sb2st_batched_f64": (_i, [_vp, _i, _l, _l, _vp, _l, _l, _vp, _l, _l, _vp, _l, _vp, _l]),
    "evd_syevd_mixed_f32": (_i, [_vp, _i, C.POINTER(_l), C.POINTER(_vp), C.POINTER(_l), C.POINTER(_vp),
                                 C.POINTER(_vp), C.POINTER(_l), C.c_char_p, _l]),
    "evd_graph_create": (_i, [_vp, _ch, _l, _l, _l, _vp, _l, _vp, _vp, _l, C.POINTER(_vp)]),
    "evd_graph_create_batched": (_i, [_vp, _ch, _i, _l, _l, _l, _vp, _l, _l, _vp, _l, _vp, _l, _l, C.POINTER(_vp)]),
    "evd_graph_create_mixed": (_i, [_vp, _i, C.POINTER(_l), C.POINTER(_vp), C.POINTER(_l), C.POINTER(_vp),
                                    C.POINTER(_vp), C.POINTER(_l), C.c_char_p, _l, C.POINTER(_vp)]),
    "evd_graph_launch": (_i, [_vp]),
    "evd_graph_wait": (_i, [_vp]),
    "evd_graph_destroy": (_i, [_vp]),
    "evd_tune_cache_put": (_i, [_l, _l, _i, C.POINTER(_i)]),
    "evd_tune_cache_get": (_i, [_l, _l, _i, C.POINTER(_i)]),
    "evd_tune_cache_clear": (_i, []),
    "evd_tune_cache_save": (_i, [C.c_char_p]),
    "evd_tune_cache_load": (_i, [C.c_char_p]),
    "evd_ctx_use_tune_cache": (_i, [_vp, _i]),
    "evd_orthogonality_f32": (_i, [_vp, _l, _l, _vp, _l, C.POINTER(_d)]),
    "evd_orthogonality_f64": (_i, [_vp, _l, _l, _vp, _l, C.POINTER(_d)]),
    "evd_backward_error_f32": (_i, [_vp, _l, _vp, _l, _vp, _vp, _l, C.POINTER(_d)]),
    "evd_sgemm": (_i, [_vp, _ch, _ch, _l, _l, _l, _f, _vp, _l, _vp, _l, _f, _vp, _l]),
    "evd_dgemm": (_i, [_vp, _ch, _ch, _l, _l, _l, _d, _vp, _l, _vp, _l, _d, _vp, _l]),
    "evd_gen_symmetric_f32": (_i, [_vp, _l, _u64, _i, _vp, _l]),
    "evd_gen_spectrum_f32": (_i, [_vp, _l, _u64, _vp, _vp, _l]),
    "evd_ctx_poison_arena": (_i, [_vp, C.c_uint]),
    "evd_server_create": (_i, [_i, _i, C.c_size_t, C.POINTER(_vp)]),
    "evd_server_set_batching": (_i, [_vp, _i]),
    "evd_server_set_precision": (_i, [_vp, _i]),
    "evd_server_set_policy": (_i, [_vp, _i]),
    "evd_server_destroy": (_i, [_vp]),
    "evd_server_submit": (_i, [_vp, _l, _l, _l, _vp, _l, _vp, _vp, _l, C.POINTER(_l)]),
    "evd_server_wait": (_i, [_vp, _l]),
    "evd_server_wait_all": (_i, [_vp]),
    "evd_server_request_ms": (_i, [_vp, _l, C.POINTER(C.c_double)]),
    "evd_server_stats": (_i, [_vp, C.POINTER(_l), C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.POINTER(C.c_double), C.POINTER(C.c_size_t)]),
}


class EvdError(RuntimeError):
    def __init__(self, func: str, code: int, msg: str):
        super().__init__(f"{func} returned {code}: {msg}")
        self.code = code
        self.func = func


def lib() -> C.CDLL:
    """Load the in-tree native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: run `python -m evd4sc2025_amd.build` (hipcc, gfx950) first")
        # The library links libamdhip64 by soname.  torch ships its own HIP
        # runtime; loading this library first would bind the process to the
        # system runtime and leave torch on a mismatched one (kernel lookups
        # then fail with hipErrorUnknown).  Importing torch first makes both
        # use torch's runtime.
        import torch  # noqa: F401

        h = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | getattr(os, "RTLD_GLOBAL", 0))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def exported_symbols() -> list[str]:
    return list(SIGNATURES)
