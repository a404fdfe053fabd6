"""ctypes binding of libwarpstate_amd.so (the C ABI in include/warpstate.h).

The library is built in-tree by ``warpstate_amd/csrc/Makefile``. There is no
CPU fallback: if the library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_char_p, c_float, c_int, c_size_t, c_uint32, c_uint64, c_void_p

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libwarpstate_amd.so")

# ws_status codes (core/src/error.rs:6-91 variants).
STATUS_NAMES = {
    0: "Ok",
    1: "EmptyPatternSet",
    2: "EmptyPattern",
    3: "NoGpuAdapter",
    4: "GpuDeviceError",
    5: "BufferMapFailed",
    6: "PatternCompilationFailed",
    7: "DfaOutOfBounds",
    8: "PatternTooLarge",
    9: "PatternSetTooLarge",
    10: "InputTooLarge",
    11: "MatchBufferOverflow",
    12: "PathologicalRegex",
    13: "InvalidArgument",
}

MODE_OVERLAPPING = 0
MODE_STANDARD = 1
MEM_HOST = 0
MEM_DEVICE = 1
VARIANT_AUTO, VARIANT_HASH, VARIANT_DFA_GLOBAL, VARIANT_DFA_LDS = 0, 1, 2, 3
VARIANT_NAMES = {VARIANT_HASH: "hash", VARIANT_DFA_GLOBAL: "dfa_global", VARIANT_DFA_LDS: "dfa_lds"}

MATCH_DTYPE = np.dtype([("pattern_id", "<u4"), ("start", "<u4"), ("end", "<u4")])
TAGGED_DTYPE = np.dtype(
    [("source_id", "<u8"), ("pattern_id", "<u4"), ("start", "<u4"), ("end", "<u4"), ("reserved", "<u4")]
)


class WsPattern(Structure):
    _fields_ = [("kind", c_uint32), ("bytes", c_void_p), ("len", c_uint64), ("name", c_char_p)]


class WsItem(Structure):
    _fields_ = [("id", c_uint64), ("data", c_void_p), ("len", c_uint64)]


class WsConfig(Structure):
    _fields_ = [
        ("max_matches", c_uint64),
        ("chunk_size", c_uint64),
        ("chunk_overlap", c_uint64),
        ("flags", c_uint32),
        ("reserved", c_uint32),
    ]


TUNING_CHECKED = 1  # ws_tuning.flags: host-checked pipeline


class WsTuning(Structure):
    _fields_ = [
        ("variant", c_uint32),
        ("bytes_per_lane", c_uint32),
        ("blocks_per_cu", c_uint32),
        ("flags", c_uint32),
    ]


class WsPatternsInfo(Structure):
    _fields_ = [
        (name, c_uint32)
        for name in (
            "pattern_count",
            "literal_count",
            "regex_count",
            "factor_count",
            "min_factor_len",
            "max_factor_len",
            "min_literal_len",
            "max_pattern_len",
            "dfa_states",
            "dfa_classes",
            "regex_dfa_states",
            "hash_window",
            "variants_supported",
            "reserved",
        )
    ]


class WsTimings(Structure):
    _fields_ = [
        ("prefilter_ms", c_float),
        ("verify_ms", c_float),
        ("regex_ms", c_float),
        ("order_ms", c_float),
        ("total_ms", c_float),
        ("reserved", c_float),
        ("hits", c_uint64),
        ("candidates", c_uint64),
        ("raw_matches", c_uint64),
    ]


INDEX_VERSION = 5
INDEX_REFERENCE_V3 = 1  # ws_index_build / ws_patterns_serialize flag: write version 3


class WsIndexInfo(Structure):
    _fields_ = [
        (name, c_uint32)
        for name in (
            "version",
            "case_insensitive",
            "pattern_count",
            "literal_count",
            "regex_count",
            "hash_window_len",
            "prebuilt",
            "reserved",
        )
    ]


# Every symbol include/warpstate.h declares, with its ctypes signature.
SIGNATURES = {
    "ws_abi_version": (c_uint32, []),
    "ws_last_error": (c_char_p, []),
    "ws_ctx_create": (c_int, [c_int, POINTER(WsConfig), POINTER(c_void_p)]),
    "ws_ctx_destroy": (None, [c_void_p]),
    "ws_ctx_stream": (c_void_p, [c_void_p]),
    "ws_ctx_cu_count": (c_uint32, [c_void_p]),
    "ws_ctx_graph_count": (c_uint32, [c_void_p]),
    "ws_ctx_set_profiling": (None, [c_void_p, c_int]),
    "ws_ctx_set_graphs": (None, [c_void_p, c_int]),
    "ws_ctx_last_timings": (c_int, [c_void_p, POINTER(WsTimings)]),
    "ws_patterns_compile": (c_int, [c_void_p, POINTER(WsPattern), c_size_t, c_uint32, POINTER(c_void_p)]),
    "ws_patterns_free": (None, [c_void_p]),
    "ws_patterns_retain": (None, [c_void_p]),
    "ws_hotswap_create": (c_int, [c_void_p, POINTER(c_void_p)]),
    "ws_hotswap_destroy": (None, [c_void_p]),
    "ws_hotswap_swap": (c_int, [c_void_p, c_void_p]),
    "ws_hotswap_scan": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_uint64, c_uint32, c_uint32, POINTER(WsTuning), c_void_p, c_uint64,
         c_uint32, POINTER(c_uint64)],
    ),
    "ws_patterns_info_get": (c_int, [c_void_p, POINTER(WsPatternsInfo)]),
    "ws_patterns_analyze": (c_int, [POINTER(WsPattern), c_size_t, c_uint32, POINTER(WsPatternsInfo)]),
    "ws_scan": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_uint64, c_uint32, c_uint32, POINTER(WsTuning), c_void_p, c_uint64,
         c_uint32, POINTER(c_uint64)],
    ),
    "ws_scan_async": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_uint64, c_uint32, c_uint32, POINTER(WsTuning), c_void_p, c_uint64,
         c_uint32],
    ),
    "ws_poll": (c_int, [c_void_p]),
    "ws_wait": (c_int, [c_void_p, POINTER(c_uint64)]),
    "ws_scan_noalloc": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_uint64, POINTER(WsTuning), c_void_p, c_uint64, POINTER(c_uint64)],
    ),
    "ws_batch_plan_create": (c_int, [c_void_p, POINTER(WsItem), c_size_t, POINTER(c_void_p)]),
    "ws_batch_plan_destroy": (None, [c_void_p]),
    "ws_scan_batch_planned": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_uint32, POINTER(WsTuning), c_void_p, c_uint64, c_uint32,
         POINTER(c_uint64)],
    ),
    "ws_scan_batch": (
        c_int,
        [c_void_p, c_void_p, POINTER(WsItem), c_size_t, c_uint32, c_uint32, POINTER(WsTuning), c_void_p,
         c_uint64, POINTER(c_uint64)],
    ),
    "ws_config_tuning": (c_int, [c_uint32, POINTER(WsTuning)]),
    "ws_router_features": (c_int, [POINTER(WsPatternsInfo), c_uint64, c_float, c_void_p]),
    "ws_autotune": (c_int, [c_void_p, c_void_p, c_void_p, c_uint64, c_uint32, c_uint32, c_uint32, c_uint32,
                            c_void_p, c_void_p, POINTER(c_uint32)]),
    "ws_router_create": (c_int, [c_int, c_uint32, POINTER(c_void_p)]),
    "ws_router_destroy": (None, [c_void_p]),
    "ws_router_last_error": (c_char_p, []),
    "ws_router_param_count": (c_uint32, []),
    "ws_router_get_params": (c_int, [c_void_p, c_void_p, c_uint32]),
    "ws_router_set_params": (c_int, [c_void_p, c_void_p, c_uint32]),
    "ws_router_set_normalization": (c_int, [c_void_p, c_void_p, c_void_p, c_float, c_float]),
    "ws_router_get_normalization": (c_int, [c_void_p, c_void_p, c_void_p, POINTER(c_float), POINTER(c_float)]),
    "ws_router_train": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_uint32, c_float,
                                c_uint32, POINTER(c_float)]),
    "ws_router_predict": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p]),
    "ws_router_forward": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p]),
    "ws_router_gradient": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_float, c_void_p]),
    "ws_stream_scan": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_uint64, c_uint32, c_uint64, c_uint64, c_uint32, POINTER(WsTuning),
         c_void_p, c_uint64, c_uint32, POINTER(c_uint64)],
    ),
    "ws_stream_scanner_create": (c_int, [c_void_p, c_void_p, c_uint32, POINTER(c_void_p)]),
    "ws_stream_scanner_destroy": (None, [c_void_p]),
    "ws_stream_scanner_feed": (c_int, [c_void_p, c_void_p, c_uint64, c_uint32, c_void_p, c_uint64,
                                       POINTER(c_uint64)]),
    "ws_stream_scanner_finish": (c_int, [c_void_p, c_void_p, c_uint64, POINTER(c_uint64)]),
    "ws_stream_scanner_processed": (c_uint64, [c_void_p]),
    "ws_index_build": (c_int, [POINTER(WsPattern), c_size_t, c_uint32, c_uint32, c_void_p, c_uint64,
                               POINTER(c_uint64)]),
    "ws_patterns_serialize": (c_int, [c_void_p, c_uint32, c_void_p, c_uint64, POINTER(c_uint64)]),
    "ws_index_inspect": (c_int, [c_void_p, c_uint64, POINTER(WsIndexInfo), POINTER(WsPatternsInfo)]),
    "ws_index_load": (c_int, [c_void_p, c_void_p, c_uint64, POINTER(c_void_p)]),
    "ws_index_load_file": (c_int, [c_void_p, c_char_p, POINTER(c_void_p)]),
    "ws_index_save_file": (c_int, [c_void_p, c_uint64, c_char_p]),
    "ws_patterns_name": (c_int, [c_void_p, c_uint32, c_void_p, c_uint64, POINTER(c_uint64)]),
}

_lib = None


class WarpstateError(Exception):
    """A ws_status error; `.variant` names the reference Error variant."""

    def __init__(self, status: int, message: str, count: int | None = None):
        self.status = status
        self.variant = STATUS_NAMES.get(status, f"Status{status}")
        self.count = count
        super().__init__(f"{self.variant}: {message}")


def load():
    """Loads the HIP library once; raises RuntimeError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"warpstate_amd HIP library not built ({LIB_PATH}); run `make -C warpstate_amd/csrc` "
            "or __graft_entry__.build()"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, count: int | None = None) -> None:
    if status != 0:
        msg = load().ws_last_error()
        raise WarpstateError(status, msg.decode(errors="replace") if msg else "", count)


def check_router(status: int) -> None:
    if status != 0:
        msg = load().ws_router_last_error()
        raise WarpstateError(status, msg.decode(errors="replace") if msg else "")
