"""ctypes binding of libtdp1375_jni.so (include/tdp1375.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
# TDP_LIB_PATH selects another build of the same library (e.g. a profiling build)
LIB_PATH = os.environ.get("TDP_LIB_PATH") or os.path.join(HERE, "libtdp1375_jni.so")
MAX_DEPTH = 15
FILTER_MAX_OUTPUT = 1 << 20

TDP_OK = 0
TDP_ERR_INVALID_ARGUMENT = 1
TDP_ERR_CANCELLED = 2
TDP_ERR_QUEUE_OVERFLOW = 3
TDP_ERR_OUT_OF_MEMORY = 4
TDP_ERR_HIP = 5
TDP_ERR_INTERNAL = 6

FLAG_DETERMINISTIC = 1
FLAG_SINGLE_WAVE = 2
FLAG_RESIDENT = 4
FLAG_GRAPH = 32
FLAG_LEARNED = 8
FLAG_SOFT_DEADLINE = 16

EXPORTED_SYMBOLS = (
    "tdp_last_error", "tdp_init_slots", "tdp_destroy_slots", "tdp_search", "tdp_search_ex",
    "tdp_search_batch", "tdp_last_launch_stats", "tdp_cancel", "tdp_filter", "tdp_eval_features",
    "tdp_filter_last_kernel_ms", "tdp_filter_noalloc", "tdp_eval_children", "tdp_host_eval_children",
    "tdp_eval_pair",
    "tdp_host_predicate_tables", "tdp_set_tuning", "tdp_set_launch_deadline", "tdp_set_expansion_batch",
    "tdp_canonical_batch",
    "tdp_set_alloc_timeout", "tdp_alloc_retries",
    "tdp_serve_start", "tdp_serve_stop", "tdp_serve_search", "tdp_serve_batch", "tdp_serve_stats",
    "tdp_device_count",
    "tdp_mlp_set_weights",
    "tdp_mlp_get_weights", "tdp_mlp_init", "tdp_mlp_forward", "tdp_mlp_gradients", "tdp_mlp_train",
    "Java_br_unb_cic_tdp_base_GPUSortingSearch_initSlots",
    "Java_br_unb_cic_tdp_base_GPUSortingSearch_destroySlots",
    "Java_br_unb_cic_tdp_base_GPUSortingSearch_searchForSortingSeq",
    "Java_br_unb_cic_tdp_base_GPUSortingSearch_cancelSort",
    "Java_br_unb_cic_tdp_base_GPUSortingSearch_filterPiPermutations",
    "Java_br_unb_cic_tdp_base_GPUSortingSearch_searchBatch",
    "Java_br_unb_cic_tdp_base_GPUSortingSearch_canonicalBatch",
)


class Counters(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "expansions", "generated", "admitted", "dominated", "improved", "inserted", "uninserted",
        "roots_live")]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


class CaseResult(ctypes.Structure):
    _fields_ = [("found", ctypes.c_int32), ("depth", ctypes.c_int32),
                ("moves", ctypes.c_int32 * (MAX_DEPTH * 3)), ("counters", Counters)]


class LaunchStats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("waves", ctypes.c_int32), ("overflowed", ctypes.c_int32),
                ("donations", ctypes.c_uint64), ("steals", ctypes.c_uint64), ("node_bytes", ctypes.c_uint64),
                ("pops", ctypes.c_uint64), ("pushes", ctypes.c_uint64), ("probes", ctypes.c_uint64),
                ("key_reads", ctypes.c_uint64), ("key_writes", ctypes.c_uint64), ("key_bytes", ctypes.c_uint64),
                ("evictions", ctypes.c_uint64), ("spills", ctypes.c_uint64),
                ("spill_takes", ctypes.c_uint64), ("overflow_info", ctypes.c_uint64), ("phase_ticks", ctypes.c_uint64 * 8)]

PHASES = ("pop", "check", "analyse", "scan", "dedup", "push", "share", "root")


class TdpError(RuntimeError):
    """A failed native call; str() is the native message (the reference's wording)."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


class CancelledError(TdpError):
    pass


class QueueOverflowError(TdpError):
    pass


_lock = threading.Lock()
_lib = None


def load() -> ctypes.CDLL:
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `python -m tdp1375_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        i32p = ctypes.POINTER(ctypes.c_int32)
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.tdp_last_error.restype = ctypes.c_char_p
        L.tdp_last_error.argtypes = []
        L.tdp_init_slots.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int]
        L.tdp_destroy_slots.argtypes = []
        L.tdp_search.argtypes = [ctypes.c_int, u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                 ctypes.c_int, i32p, i32p]
        L.tdp_search_ex.argtypes = [ctypes.c_int, u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_uint32, i32p, i32p, ctypes.POINTER(Counters)]
        L.tdp_search_batch.argtypes = [ctypes.c_int, ctypes.c_int, u8p, u8p, i32p, i32p,
                                       ctypes.POINTER(ctypes.c_float), i32p, i32p, ctypes.c_uint32,
                                       ctypes.POINTER(CaseResult)]
        L.tdp_last_launch_stats.argtypes = [ctypes.c_int, ctypes.POINTER(LaunchStats)]
        L.tdp_cancel.argtypes = [ctypes.c_int]
        L.tdp_cancel.restype = None
        L.tdp_filter.argtypes = [u8p, ctypes.c_int, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, u8p,
                                 ctypes.c_int64, i64p, i64p]
        L.tdp_filter_last_kernel_ms.restype = ctypes.c_double
        L.tdp_filter_last_kernel_ms.argtypes = []
        L.tdp_filter_noalloc.argtypes = [u8p, ctypes.c_int, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L.tdp_set_alloc_timeout.argtypes = [ctypes.c_int64]
        L.tdp_serve_start.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                      ctypes.c_int]
        L.tdp_serve_stop.argtypes = []
        L.tdp_serve_search.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                       ctypes.c_int, i32p, i32p, ctypes.POINTER(Counters)]
        L.tdp_serve_batch.argtypes = [ctypes.c_int, u8p, u8p, i32p, i32p, ctypes.POINTER(ctypes.c_float), i32p, i32p,
                                      ctypes.POINTER(CaseResult)]
        L.tdp_serve_stats.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_int64)]
        L.tdp_alloc_retries.argtypes = []
        L.tdp_alloc_retries.restype = ctypes.c_uint64
        L.tdp_eval_children.argtypes = [u8p, u8p, ctypes.c_int, i32p]
        L.tdp_eval_pair.argtypes = [u8p, u8p, u8p, u8p] + [ctypes.c_int] * 3 + [ctypes.c_uint32] * 4 + [i32p]
        L.tdp_host_eval_children.argtypes = [u8p, u8p, ctypes.c_int, i32p]
        L.tdp_eval_features.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_uint32)]
        L.tdp_host_predicate_tables.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                                ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                                ctypes.POINTER(ctypes.c_uint64)]
        L.tdp_set_tuning.argtypes = [ctypes.c_int] * 5
        L.tdp_set_launch_deadline.argtypes = [ctypes.c_int]
        L.tdp_set_expansion_batch.argtypes = [ctypes.c_int]
        L.tdp_canonical_batch.argtypes = [ctypes.c_int, u8p, u8p, i32p, ctypes.POINTER(ctypes.c_float), i32p, i32p,
                                          u8p]
        L.tdp_device_count.argtypes = []
        f32p = ctypes.POINTER(ctypes.c_float)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.tdp_mlp_set_weights.argtypes = [f32p, f32p, f32p, ctypes.c_float]
        L.tdp_mlp_get_weights.argtypes = [f32p, f32p, f32p, f32p]
        L.tdp_mlp_init.argtypes = [ctypes.c_uint64]
        L.tdp_mlp_forward.argtypes = [u32p, ctypes.c_int, f32p]
        L.tdp_mlp_gradients.argtypes = [u32p, u8p, i32p, ctypes.c_int, f32p, f32p]
        L.tdp_mlp_train.argtypes = [u32p, u8p, i32p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                    ctypes.c_uint64, f32p]
        _lib = L
        return L


def check(status: int) -> None:
    if status == TDP_OK:
        return
    msg = load().tdp_last_error().decode(errors="replace")
    if status == TDP_ERR_CANCELLED:
        raise CancelledError(status, msg)
    if status == TDP_ERR_QUEUE_OVERFLOW:
        raise QueueOverflowError(status, msg)
    raise TdpError(status, msg)


def u8(seq, size=None):
    n = size if size is not None else max(1, len(seq))
    arr = (ctypes.c_uint8 * n)()
    for i, v in enumerate(seq):
        arr[i] = int(v) & 0xFF
    return arr
