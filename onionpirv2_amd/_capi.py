"""ctypes binding of include/onionpir_gpu.h (libonionpir.so).

The library must have been built (``python -m onionpirv2_amd.build``); loading
fails loudly otherwise -- there is no CPU fallback for the server path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "_lib", "libonionpir.so")

U64P = C.POINTER(C.c_uint64)

ONION_OK = 0
ERRORS = {-1: "EINVAL", -2: "ENOKEY", -3: "ENOMEM", -4: "EHIP", -5: "ESTATE"}


class OnionParams(C.Structure):
    _fields_ = [("poly_degree", C.c_uint32), ("plain_bits", C.c_uint32), ("gsw_l", C.c_uint32),
                ("gsw_l_key", C.c_uint32), ("fst_dim", C.c_uint32), ("other_dim", C.c_uint32)]


class OnionParamInfo(C.Structure):
    _fields_ = [("q", C.c_uint64), ("p", C.c_uint64), ("t", C.c_uint64), ("small_q", C.c_uint64),
                ("base_log2", C.c_uint32), ("base_log2_key", C.c_uint32), ("ndim", C.c_uint32),
                ("expan_height", C.c_uint32), ("useful_cnt", C.c_uint32), ("num_pt", C.c_uint64),
                ("galois_key_words", C.c_uint32), ("rgsw_key_words", C.c_uint32),
                ("response_bytes", C.c_uint32), ("poly_degree", C.c_uint32)]


class OnionStageTimes(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("expand", "convert", "first_dim", "other_dim", "mod_switch", "total", "first_dim_core")]


class OnionTuning(C.Structure):
    _fields_ = [("fuse_min", C.c_uint32), ("fd_rows_b1", C.c_uint32), ("fd_tile_rows", C.c_uint32),
                ("fd_b1_variant", C.c_uint32), ("use_graph", C.c_uint32), ("fd_mfma_min", C.c_uint32),
                ("fd_mfma_variant", C.c_uint32), ("db_planes", C.c_uint32),
                ("ntt_radix", C.c_uint32), ("batch_split", C.c_uint32)]


class OnionError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class OnionKeyError(OnionError, KeyError):
    """Unknown client id (the reference raises std::out_of_range)."""


class OnionValueError(OnionError, ValueError):
    """Invalid argument (the reference raises std::invalid_argument)."""


_lib = None

_SIGS = {
    "onion_gpu_last_error": (C.c_char_p, []),
    "onion_gpu_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "onion_params_derive": (C.c_int, [C.POINTER(OnionParams), C.POINTER(OnionParamInfo)]),
    "onion_gpu_create": (C.c_int, [C.POINTER(OnionParams), C.POINTER(C.c_void_p)]),
    "onion_gpu_destroy": (C.c_int, [C.c_void_p]),
    "onion_gpu_param_info": (C.c_int, [C.c_void_p, C.POINTER(OnionParamInfo)]),
    "onion_gpu_db_generate": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int]),
    "onion_gpu_db_upload_plain": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, U64P]),
    "onion_gpu_db_read_ntt": (C.c_int, [C.c_void_p, C.c_uint64, U64P]),
    "onion_gpu_set_galois_key": (C.c_int, [C.c_void_p, C.c_uint64, U64P, C.c_size_t]),
    "onion_gpu_set_rgsw_key": (C.c_int, [C.c_void_p, C.c_uint64, U64P, C.c_size_t]),
    "onion_gpu_drop_client": (C.c_int, [C.c_void_p, C.c_uint64]),
    "onion_gpu_share_keys": (C.c_int, [C.c_void_p, C.c_void_p]),
    "onion_gpu_answer": (C.c_int, [C.c_void_p, C.c_uint64, U64P, U64P]),
    "onion_gpu_answer_batch": (C.c_int, [C.c_void_p, C.c_uint32, U64P, U64P, U64P]),
    "onion_gpu_pack_response": (C.c_int, [C.c_void_p, U64P, C.c_char_p, C.c_size_t,
                                          C.POINTER(C.c_size_t)]),
    "onion_gpu_last_stage_times": (C.c_int, [C.c_void_p, C.POINTER(OnionStageTimes)]),
    "onion_gpu_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(OnionTuning)]),
    "onion_gpu_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(OnionTuning)]),
    "onion_gpu_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32,
                                     C.POINTER(OnionTuning)]),
    "onion_gpu_set_lanes": (C.c_int, [C.c_void_p, C.c_uint32]),
    "onion_gpu_answer_batch_lane": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, U64P, U64P, U64P,
                                              C.POINTER(OnionStageTimes)]),
    "onion_gpu_answer_batch_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, U64P, U64P, U64P]),
    "onion_gpu_scratch_bytes_per_query": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "onion_gpu_scratch_bytes": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint64)]),
    "onion_gpu_reserve": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "onion_gpu_tap_expand": (C.c_int, [C.c_void_p, C.c_uint64, U64P, U64P]),
    "onion_gpu_tap_convert": (C.c_int, [C.c_void_p, C.c_uint64, U64P, U64P]),
    "onion_gpu_tap_first_dim": (C.c_int, [C.c_void_p, U64P, U64P]),
    "onion_gpu_tap_first_dim_batch": (C.c_int, [C.c_void_p, C.c_uint32, U64P, U64P]),
    "onion_gpu_tap_mux": (C.c_int, [C.c_void_p, U64P, C.c_uint32, U64P, U64P]),
    "onion_gpu_tap_ntt": (C.c_int, [C.c_void_p, C.c_int, C.c_int, U64P, C.c_uint32]),
    "onion_gpu_tap_galois": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, U64P, U64P]),
    "onion_gpu_tap_mod_switch": (C.c_int, [C.c_void_p, U64P, U64P]),
    "onion_gpu_tap_copy": (C.c_int, [C.c_void_p, C.c_uint64, U64P]),
    "onion_gpu_tap_valu_probe": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_double)]),
}


def lib():
    """Load libonionpir.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m onionpirv2_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def device_count() -> int:
    n = C.c_int(0)
    check(lib().onion_gpu_device_count(C.byref(n)))
    return n.value


def exported_symbols():
    return list(_SIGS)


def check(rc: int) -> None:
    if rc != ONION_OK:
        msg = lib().onion_gpu_last_error().decode(errors="replace")
        if rc == -2:
            raise OnionKeyError(rc, msg)
        if rc == -1:
            raise OnionValueError(rc, msg)
        raise OnionError(rc, msg)


def u64ptr(a: np.ndarray):
    if a.dtype != np.uint64 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous uint64 array")
    return a.ctypes.data_as(U64P)
