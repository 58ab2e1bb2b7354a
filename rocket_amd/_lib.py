"""ctypes binding of libprnu_hip.so (the C ABI declared in include/prnu_api.h).

Loading fails loudly: there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "libprnu_hip.so")
HEADER = os.path.join(ROOT, "include", "prnu_api.h")

# status codes (prnu_status)
OK, EINVAL, ENOMEM, EDEVICE, ENOTSUP = 0, 1, 2, 3, 4
MODE_NCC, MODE_PCE = 0, 1
INPUT_RGB8, INPUT_GRAY8, INPUT_FLOAT32 = 0, 1, 2
PREC_FP32, PREC_SPLIT_BF16, PREC_FP8, PREC_FP16 = 0, 1, 2, 3
PCE_REFERENCE, PCE_TEXTBOOK = 0, 1
INDEX_HILBERT, INDEX_TILING = 0, 1
PLAN_PER_ITEM, PLAN_PERSISTENT = 0, 1
JPEG_CORRUPT_DATA = 16

PRECISIONS = {"fp32": PREC_FP32, "split_bf16": PREC_SPLIT_BF16, "fp8": PREC_FP8, "fp16": PREC_FP16}


class PrnuError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[status {status}] {msg}")
        self.status = status


class PrnuInvalidArgument(PrnuError, ValueError):
    """EINVAL: the reference raises IllegalArgumentException here."""


_c_void_p = ctypes.c_void_p
_c_int = ctypes.c_int
_c_i64 = ctypes.c_int64
_c_size = ctypes.c_size_t
_c_dbl_p = ctypes.c_void_p

_SIGNATURES = {
    "prnu_last_error": (ctypes.c_char_p, []),
    "prnu_api_version": (_c_int, []),
    "prnu_device_count": (_c_int, [ctypes.POINTER(_c_int)]),
    "prnu_set_device": (_c_int, [_c_int]),
    "prnu_mem_info": (_c_int, [ctypes.POINTER(_c_size), ctypes.POINTER(_c_size)]),
    "prnu_malloc_device": (_c_int, [ctypes.POINTER(_c_void_p), _c_size]),
    "prnu_free_device": (_c_int, [_c_void_p]),
    "prnu_malloc_pinned": (_c_int, [ctypes.POINTER(_c_void_p), _c_size]),
    "prnu_free_pinned": (_c_int, [_c_void_p]),
    "prnu_memcpy_h2d": (_c_int, [_c_void_p, _c_void_p, _c_size, _c_void_p]),
    "prnu_memcpy_d2h": (_c_int, [_c_void_p, _c_void_p, _c_size, _c_void_p]),
    "prnu_memcpy_d2d": (_c_int, [_c_void_p, _c_void_p, _c_size, _c_void_p]),
    "prnu_stream_create": (_c_int, [ctypes.POINTER(_c_void_p)]),
    "prnu_stream_destroy": (_c_int, [_c_void_p]),
    "prnu_stream_synchronize": (_c_int, [_c_void_p]),
    "prnu_ctx_create": (_c_int, [_c_int, _c_int, _c_int, ctypes.POINTER(_c_void_p)]),
    "prnu_ctx_destroy": (_c_int, [_c_void_p]),
    "prnu_max_file_size": (_c_i64, [_c_void_p]),
    "prnu_max_input_size": (_c_i64, [_c_void_p]),
    "prnu_max_output_size": (_c_i64, [_c_void_p]),
    "prnu_preprocess": (_c_int, [_c_void_p, _c_int, _c_void_p, _c_size, _c_void_p, _c_size,
                                 ctypes.POINTER(_c_size), _c_void_p]),
    "prnu_preprocess_batch": (_c_int, [_c_void_p, _c_int, _c_void_p, _c_size, _c_i64, _c_int, _c_void_p, _c_size,
                                       _c_i64, _c_void_p]),
    "prnu_correlate": (_c_int, [_c_void_p, _c_void_p, _c_size, _c_void_p, _c_size, _c_void_p, _c_size,
                                ctypes.POINTER(_c_size), _c_void_p]),
    "prnu_correlate_pce": (_c_int, [_c_void_p, _c_int, _c_void_p, _c_void_p, _c_void_p, _c_void_p]),
    "prnu_pce_pairs": (_c_int, [_c_void_p, _c_int, _c_void_p, _c_void_p, _c_int, _c_void_p, _c_void_p]),
    "prnu_filter_grayscale": (_c_int, [_c_void_p, _c_int, _c_void_p, _c_size, _c_void_p, _c_void_p]),
    "prnu_filter_fastnoise": (_c_int, [_c_void_p, _c_void_p, _c_void_p]),
    "prnu_filter_zeromean": (_c_int, [_c_void_p, _c_void_p, _c_void_p]),
    "prnu_filter_wiener": (_c_int, [_c_void_p, _c_void_p, _c_void_p]),
    "prnu_filter_spectral": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_size, _c_void_p]),
    "prnu_pattern_stats": (_c_int, [_c_void_p, _c_int, _c_i64, _c_void_p, _c_void_p, _c_void_p, _c_void_p]),
    "prnu_measure_mfma_peak": (_c_int, [_c_int, _c_int, _c_int, _c_void_p, _c_void_p]),
    "prnu_pattern_stats_noalloc": (_c_int, [_c_void_p, _c_int, _c_i64, _c_void_p, _c_void_p, _c_void_p, _c_void_p]),
    "prnu_store_create": (_c_int, [_c_i64, _c_int, _c_int, ctypes.POINTER(_c_void_p)]),
    "prnu_store_destroy": (_c_int, [_c_void_p]),
    "prnu_store_info": (_c_int, [_c_void_p, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_int), ctypes.POINTER(_c_int),
                                 ctypes.POINTER(_c_size)]),
    "prnu_store_put": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_void_p]),
    "prnu_store_stats": (_c_int, [_c_void_p, ctypes.POINTER(_c_void_p), ctypes.POINTER(_c_void_p),
                                  ctypes.POINTER(_c_void_p)]),
    "prnu_store_gram_ncc": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_void_p, _c_void_p]),
    "prnu_gram_plan_create": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_i64, ctypes.POINTER(_c_void_p)]),
    "prnu_gram_plan_create_blocks": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_i64, ctypes.POINTER(_c_void_p)]),
    "prnu_gram_plan_destroy": (_c_int, [_c_void_p]),
    "prnu_gram_plan_run": (_c_int, [_c_void_p, _c_int, _c_void_p]),
    "prnu_gram_plan_autotune": (_c_int, [_c_void_p, _c_void_p, ctypes.POINTER(_c_i64), ctypes.POINTER(ctypes.c_float)]),
    "prnu_gram_plan_set_launch": (_c_int, [_c_void_p, _c_int]),
    "prnu_gram_plan_get_launch": (_c_int, [_c_void_p, ctypes.POINTER(_c_int)]),
    "prnu_gram_plan_info": (_c_int, [_c_void_p, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_int), ctypes.POINTER(_c_i64),
                                     ctypes.POINTER(_c_size), ctypes.POINTER(ctypes.c_double)]),
    "prnu_fpstore_create": (_c_int, [_c_i64, _c_int, ctypes.POINTER(_c_void_p)]),
    "prnu_fpstore_destroy": (_c_int, [_c_void_p]),
    "prnu_fpstore_info": (_c_int, [_c_void_p, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_int),
                                   ctypes.POINTER(_c_int), ctypes.POINTER(_c_size)]),
    "prnu_fpstore_put": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_void_p]),
    "prnu_fpstore_query": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_int, _c_int, _c_void_p, _c_void_p,
                                    _c_void_p, _c_void_p]),
    "prnu_gram_set_config": (_c_int, [_c_void_p, _c_int, _c_void_p]),
    "prnu_gram_get_config": (_c_int, [_c_void_p, _c_int, _c_int, _c_void_p]),
    "prnu_gram_autotune": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_int, _c_void_p,
                                    ctypes.POINTER(ctypes.c_float), _c_void_p]),
    "prnu_triangle_size": (_c_i64, [_c_int, _c_int, _c_int, _c_int, _c_int]),
    "prnu_index_divide": (_c_int, [_c_int, _c_void_p, _c_int, _c_int, _c_void_p, _c_int, ctypes.POINTER(_c_int),
                                   _c_void_p, _c_i64, ctypes.POINTER(_c_i64)]),
    "prnu_index_pairs": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_void_p, _c_i64, ctypes.POINTER(_c_i64)]),
    "prnu_slotcache_slots_for": (_c_int, [_c_size, _c_i64, _c_size, ctypes.POINTER(_c_i64)]),
    "prnu_slotcache_create": (_c_int, [_c_i64, ctypes.POINTER(_c_void_p)]),
    "prnu_slotcache_destroy": (_c_int, [_c_void_p]),
    "prnu_slotcache_acquire": (_c_int, [_c_void_p, _c_void_p, _c_int, ctypes.POINTER(_c_i64)]),
    "prnu_slotcache_poll": (_c_int, [_c_void_p, _c_i64, ctypes.POINTER(_c_int), _c_void_p, _c_void_p, _c_void_p,
                                     _c_int, ctypes.POINTER(_c_int)]),
    "prnu_slotcache_ready": (_c_int, [_c_void_p, _c_i64, _c_i64]),
    "prnu_slotcache_release": (_c_int, [_c_void_p, _c_i64]),
    "prnu_slotcache_try_acquire": (_c_int, [_c_void_p, _c_i64, ctypes.POINTER(_c_i64)]),
    "prnu_slotcache_meta": (_c_int, [_c_void_p, _c_i64, ctypes.POINTER(_c_i64)]),
    "prnu_slotcache_entry": (_c_int, [_c_void_p, _c_i64, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_int),
                                      ctypes.POINTER(_c_i64)]),
    "prnu_slotcache_lru": (_c_int, [_c_void_p, _c_void_p, _c_i64, ctypes.POINTER(_c_i64)]),
    "prnu_slotcache_stats": (_c_int, [_c_void_p, _c_void_p]),
    "prnu_slotcache_cleanup": (_c_int, [_c_void_p]),
    "prnu_jpeg_decoder_create": (_c_int, [ctypes.POINTER(_c_void_p)]),
    "prnu_jpeg_decoder_destroy": (_c_int, [_c_void_p]),
    "prnu_jpeg_info": (_c_int, [_c_void_p, _c_i64, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "prnu_jpeg_decode_batch": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_int, _c_void_p, _c_i64, _c_i64,
                                        _c_void_p, _c_void_p, _c_void_p, _c_void_p]),
    "readJPEG": (_c_int, [_c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                          _c_void_p, ctypes.c_int32]),
    "prnu_store_image_bytes": (_c_i64, [_c_void_p, _c_int]),
    "prnu_store_save": (_c_int, [_c_void_p, _c_int, _c_int, _c_void_p, _c_void_p]),
    "prnu_store_load": (_c_int, [_c_void_p, _c_int, _c_void_p, _c_size, _c_void_p]),
    "prnu_ctx_set_timing": (_c_int, [_c_void_p, _c_int]),
    "prnu_ctx_last_kernel_ms": (_c_int, [_c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "prnu_gram_ncc": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_void_p, _c_void_p, _c_void_p,
                               _c_void_p, _c_void_p]),
}

_lib = None


def header_symbols() -> list[str]:
    """Every function name include/prnu_api.h declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"PRNU_EXPORT\s+[\w\s\*]*?\b(prnu_\w+|readJPEG)\s*\(", text)))


def load() -> ctypes.CDLL:
    """Load the HIP library.  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m rocket_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, what: str = "") -> None:
    if status == OK:
        return
    msg = load().prnu_last_error().decode(errors="replace")
    if what:
        msg = f"{what}: {msg}"
    if status == EINVAL:
        raise PrnuInvalidArgument(status, msg)
    raise PrnuError(status, msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)
