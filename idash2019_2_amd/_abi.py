"""ctypes declarations of the C ABI in include/idash_hip.h (libidash_hip.so).

The library is the product: there is no CPU fallback. If the shared object is
missing, importing this module raises, so a missing build fails loudly.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libidash_hip.so")

IHE_OK = 0
IHE_ERR_ARG = 1
IHE_ERR_HIP = 2
IHE_ERR_RANGE = 3
IHE_ERR_NOMEM = 4
IHE_ERR_UNSUPPORTED = 5
IHE_ERR_BUSY = 6

IHE_N = 1024
IHE_CONSTANT_BIDX = 0xFFFFFFFF
IHE_NOISE_TABLE_LEN = 1280

VARIANT_AUTO, VARIANT_VALU, VARIANT_WIDE, VARIANT_MFMA = 0, 1, 2, 3
(TIMER_CLOUD, TIMER_CLOUD_VAR, TIMER_ENCRYPT, TIMER_DECRYPT, TIMER_BATCH, TIMER_STREAM,
 TIMER_REGION) = range(7)

# Every entry point the header declares, with (restype, argtypes).
_vp, _u32, _i32, _u64, _sz, _int = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32,
                                    ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int)
_pvp = ctypes.POINTER(ctypes.c_void_p)


class Layout(ctypes.Structure):
    _fields_ = [("num_samples", _u32), ("num_regions", _u32), ("region_size", _u32)]


class ModelCSR(ctypes.Structure):
    _fields_ = [("num_outputs", _u32), ("row_ptr", _vp), ("in_bidx", _vp), ("weight", _vp)]


class Cohort(ctypes.Structure):
    _fields_ = [("band", _vp), ("layout", Layout), ("num_in_cts", _u32), ("d_in", _vp),
                ("d_in_var", _vp), ("d_out", _vp), ("d_out_var", _vp)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("num_cohorts", _u32), ("num_launches", _u32), ("num_tiles", _u32),
                ("captured", _u32), ("persistent", _u32)]


class TuneEntry(ctypes.Structure):
    _fields_ = [("rows_per_tile", _u32), ("kblocks", _u32), ("col_tiles_per_wave", _u32),
                ("ms", ctypes.c_float)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", _u64), ("num_slots", _u32), ("queue_capacity", _u32)]


class Request(ctypes.Structure):
    _fields_ = [("model", _u32), ("num_samples", _u32), ("geno", _vp), ("key", _vp),
                ("seed", _u64), ("out", _vp), ("arrival_ns", _u64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", _u64), ("pending", _u32), ("running", _u32),
                ("arena_bytes", _u64), ("arena_used", _u64), ("arena_peak", _u64)]


class BandInfo(ctypes.Structure):
    _fields_ = [("num_outputs", _u32), ("num_groups", _u32), ("num_tiles", _u32),
                ("num_cols", _u32), ("nnz", _u32), ("variant", _u32), ("max_window", _u32),
                ("device_bytes", _u64), ("tile_rows", _u32), ("tile_kblocks", _u32),
                ("col_tiles", _u32)]


SIGNATURES = {
    "ihe_last_error": (ctypes.c_char_p, []),
    "ihe_device_count": (_int, [ctypes.POINTER(_int)]),
    "ihe_ctx_create": (_int, [_int, _pvp]),
    "ihe_ctx_destroy": (_int, [_vp]),
    "ihe_ctx_set_stream": (_int, [_vp, _vp]),
    "ihe_ctx_get_stream": (_int, [_vp, _pvp]),
    "ihe_sync": (_int, [_vp]),
    "ihe_malloc": (_int, [_vp, _sz, _pvp]),
    "ihe_free": (_int, [_vp, _vp]),
    "ihe_host_alloc": (_int, [_vp, _sz, _pvp]),
    "ihe_host_free": (_int, [_vp, _vp]),
    "ihe_memcpy_h2d": (_int, [_vp, _vp, _vp, _sz]),
    "ihe_memcpy_d2h": (_int, [_vp, _vp, _vp, _sz]),
    "ihe_memcpy_h2d_async": (_int, [_vp, _vp, _vp, _sz]),
    "ihe_memcpy_d2h_async": (_int, [_vp, _vp, _vp, _sz]),
    "ihe_memset": (_int, [_vp, _vp, _int, _sz]),
    "ihe_memcpy_d2d": (_int, [_vp, _vp, _vp, _sz]),
    "ihe_checksum": (_int, [_vp, _vp, _sz, ctypes.POINTER(_u64)]),
    "ihe_stream_probe": (_int, [_vp, _vp, _sz, _vp, _sz]),
    "ihe_stream_probe_ex": (_int, [_vp, _vp, _sz, _vp, _sz, _u32, _u32, _u32]),
    "ihe_timing_enable": (_int, [_vp, _int]),
    "ihe_timing_reset": (_int, [_vp]),
    "ihe_timing_region": (_int, [_vp, _int]),
    "ihe_timing_read": (_int, [_vp, _int, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(_u32)]),
    "ihe_band_create": (_int, [_vp, ctypes.POINTER(Layout), _u32, ctypes.POINTER(ModelCSR), _int,
                               _pvp]),
    "ihe_band_destroy": (_int, [_vp]),
    "ihe_band_get_info": (_int, [_vp, ctypes.POINTER(BandInfo)]),
    "ihe_band_plan": (_int, [ctypes.POINTER(Layout), _u32, ctypes.POINTER(ModelCSR), _int,
                             ctypes.POINTER(BandInfo)]),
    "ihe_cloud_score": (_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "ihe_cloud_score_layout": (_int, [_vp, _vp, ctypes.POINTER(Layout), _u32, _vp, _vp, _vp, _vp]),
    "ihe_band_create_tuned": (_int, [_vp, ctypes.POINTER(Layout), _u32, ctypes.POINTER(ModelCSR),
                                     _vp, _vp, _u32, _pvp, ctypes.POINTER(TuneEntry),
                                     ctypes.POINTER(_u32)]),
    "ihe_batch_create": (_int, [_vp, ctypes.POINTER(Cohort), _u32, _pvp]),
    "ihe_batch_capture": (_int, [_vp]),
    "ihe_batch_run": (_int, [_vp]),
    "ihe_batch_get_info": (_int, [_vp, ctypes.POINTER(BatchInfo)]),
    "ihe_batch_destroy": (_int, [_vp]),
    "ihe_server_create": (_int, [_vp, ctypes.POINTER(ServerConfig), _pvp]),
    "ihe_server_add_model": (_int, [_vp, _vp, _u32, ctypes.POINTER(_u32)]),
    "ihe_server_submit": (_int, [_vp, ctypes.POINTER(Request), ctypes.POINTER(_u64)]),
    "ihe_server_poll": (_int, [_vp, _u64, ctypes.POINTER(_int), ctypes.POINTER(ctypes.c_double)]),
    "ihe_server_wait": (_int, [_vp, _u64, ctypes.POINTER(ctypes.c_double)]),
    "ihe_server_get_stats": (_int, [_vp, ctypes.POINTER(ServerStats)]),
    "ihe_server_destroy": (_int, [_vp]),
    "ihe_key_create": (_int, [_vp, _vp, _pvp]),
    "ihe_key_destroy": (_int, [_vp]),
    "ihe_encrypt": (_int, [_vp, _vp, ctypes.POINTER(Layout), _u32, _vp, _u64, _vp, _vp]),
    "ihe_decrypt": (_int, [_vp, _vp, _u32, _vp, _u32, _vp]),
    "ihe_decrypt_noalloc": (_int, [_vp, _u32, _vp, _u32, _vp, _vp]),
    "ihe_noise_table": (_int, [_vp, _u32]),
}

_LIB = None


class IheError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[ihe error {code}] {msg}")
        self.code = code


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libidash_hip.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc: int) -> None:
    if rc != IHE_OK:
        msg = lib().ihe_last_error()
        raise IheError(rc, msg.decode() if msg else "")
