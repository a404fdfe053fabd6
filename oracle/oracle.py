"""ctypes wrapper of the C oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker; the product package never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, Structure, c_char_p, c_int, c_int64, c_size_t, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
MATCH_DTYPE = np.dtype([("pattern_id", "<u4"), ("start", "<u4"), ("end", "<u4")])


class OrPattern(Structure):
    _fields_ = [("kind", c_uint32), ("bytes", c_void_p), ("len", c_uint64)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.or_fnv1a_window.restype = c_uint32
        L.or_fnv1a_window.argtypes = [c_void_p, c_uint64, c_uint32]
        L.or_prefilter_sizes.restype = None
        L.or_prefilter_sizes.argtypes = [c_void_p, c_uint32, c_uint32, POINTER(c_uint32), POINTER(c_uint32)]
        L.or_prefilter_build.restype = c_int
        L.or_prefilter_build.argtypes = [c_void_p, c_void_p, c_uint32, c_uint32, c_void_p, c_void_p, c_void_p]
        L.or_prefilter_probe.restype = c_uint32
        L.or_prefilter_probe.argtypes = [c_void_p, c_uint32, c_void_p, c_uint32, c_void_p, c_uint32, c_uint32,
                                         c_void_p, c_uint32]
        L.or_compile.restype = c_void_p
        L.or_compile.argtypes = [POINTER(OrPattern), c_uint32, c_int, POINTER(c_int), c_char_p, c_size_t]
        L.or_free.restype = None
        L.or_free.argtypes = [c_void_p]
        for name in ("or_literal_count", "or_regex_count", "or_strategy"):
            getattr(L, name).restype = c_uint32
            getattr(L, name).argtypes = [c_void_p]
        L.or_scan.restype = c_uint64
        L.or_scan.argtypes = [c_void_p, c_void_p, c_uint64, c_int, c_void_p, c_uint64]
        L.or_scan_overlapping_mt.restype = c_uint64
        L.or_scan_overlapping_mt.argtypes = [c_void_p, c_void_p, c_uint64, c_uint32, c_void_p, c_uint64]
        L.or_scan_mt.restype = c_uint64
        L.or_scan_mt.argtypes = [c_void_p, c_void_p, c_uint64, c_int, c_uint32, c_void_p, c_uint64]
        L.or_regex_find_iter.restype = c_uint64
        L.or_regex_find_iter.argtypes = [c_void_p, c_uint32, c_void_p, c_uint64, c_void_p, c_uint64]
        L.or_regex_match_at.restype = c_int64
        L.or_regex_match_at.argtypes = [c_void_p, c_uint32, c_void_p, c_uint64, c_uint64]
        _lib = L
    return _lib


def _b(x) -> bytes:
    return x.encode() if isinstance(x, str) else bytes(x)


def fnv1a_window(data, window: int) -> int:
    b = _b(data)
    return int(lib().or_fnv1a_window(b, len(b), window))


def prefilter_table(lens, hashes, window: int):
    """LiteralPrefilterTable::build as numpy arrays (prefix_meta, ranges, entries)."""
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    hashes = np.ascontiguousarray(hashes, dtype=np.uint32)
    nb, ne = c_uint32(), c_uint32()
    lib().or_prefilter_sizes(lens.ctypes.data, len(lens), window, ctypes.byref(nb), ctypes.byref(ne))
    meta = np.zeros(4 * window, dtype=np.uint32)
    ranges = np.zeros(2 * max(nb.value, 1), dtype=np.uint32)
    entries = np.zeros(2 * max(ne.value, 1), dtype=np.uint32)
    lib().or_prefilter_build(lens.ctypes.data, hashes.ctypes.data, len(lens), window, meta.ctypes.data,
                             ranges.ctypes.data, entries.ctypes.data)
    return meta, ranges[: 2 * nb.value], entries[: 2 * ne.value]


def prefilter_probe(meta, ranges, entries, window: int, prefix_len: int, h: int):
    out = np.zeros(1024, dtype=np.uint32)
    n = lib().or_prefilter_probe(meta.ctypes.data, window, ranges.ctypes.data, len(ranges) // 2,
                                 entries.ctypes.data, prefix_len, h, out.ctypes.data, len(out))
    return out[: min(n, len(out))].tolist()


class OracleError(Exception):
    def __init__(self, status: int, msg: str):
        self.status = status
        super().__init__(f"status {status}: {msg}")


class Oracle:
    """A compiled pattern set on the CPU. patterns: list of (kind, bytes|str)."""

    def __init__(self, patterns, case_insensitive: bool = False):
        self._keep = [_b(p) for _, p in patterns]
        arr = (OrPattern * len(patterns))()
        for i, ((kind, _), b) in enumerate(zip(patterns, self._keep)):
            arr[i].kind = kind
            arr[i].bytes = ctypes.cast(ctypes.c_char_p(b), c_void_p) if b else None
            arr[i].len = len(b)
        st = c_int(0)
        err = ctypes.create_string_buffer(256)
        h = lib().or_compile(arr, len(patterns), 1 if case_insensitive else 0, ctypes.byref(st), err, 256)
        if not h:
            raise OracleError(st.value, err.value.decode(errors="replace"))
        self.h = c_void_p(h)

    def __del__(self):
        try:
            if self.h:
                lib().or_free(self.h)
        except Exception:
            pass

    @property
    def strategy(self) -> int:
        return int(lib().or_strategy(self.h))

    def scan(self, data, mode: int = 0) -> np.ndarray:
        """mode 0: scan_overlapping; mode 1: PatternSet::scan."""
        buf = np.frombuffer(_b(data) if not isinstance(data, np.ndarray) else data, dtype=np.uint8)
        n = lib().or_scan(self.h, buf.ctypes.data, len(buf), mode, None, 0)
        out = np.zeros(max(int(n), 1), dtype=MATCH_DTYPE)
        lib().or_scan(self.h, buf.ctypes.data, len(buf), mode, out.ctypes.data, n)
        return out[: int(n)]

    def scan_into(self, data, out: np.ndarray, mode: int = 0, threads: int = 1) -> int:
        """One pass into a caller buffer (MATCH_DTYPE); returns the total count.
        threads > 1 splits an overlapping literal-only scan by byte range."""
        buf = np.frombuffer(_b(data) if not isinstance(data, np.ndarray) else data, dtype=np.uint8)
        if threads > 1:
            if mode != 0:
                raise ValueError("the threaded oracle scan is overlapping-only")
            n = lib().or_scan_overlapping_mt(self.h, buf.ctypes.data, len(buf), threads, out.ctypes.data, len(out))
            if n == 2 ** 64 - 1:
                raise ValueError("the threaded oracle scan is literal-only")
            return int(n)
        return int(lib().or_scan(self.h, buf.ctypes.data, len(buf), mode, out.ctypes.data, len(out)))

    def scan_mt(self, data, mode: int = 0, threads: int = 16) -> np.ndarray:
        """scan() over `threads` threads (literal shards and one job per regex)."""
        buf = np.frombuffer(_b(data) if not isinstance(data, np.ndarray) else data, dtype=np.uint8)
        cap = max(1 << 16, len(buf) // 64)
        while True:
            out = np.zeros(cap, dtype=MATCH_DTYPE)
            n = int(lib().or_scan_mt(self.h, buf.ctypes.data, len(buf), mode, threads, out.ctypes.data, cap))
            if n <= cap:
                return out[:n]
            cap = n

    def scan_list(self, data, mode: int = 0):
        a = self.scan(data, mode)
        return [(int(p), int(s), int(e)) for p, s, e in zip(a["pattern_id"], a["start"], a["end"])]

    def find_iter(self, regex_index: int, data):
        buf = np.frombuffer(_b(data), dtype=np.uint8)
        n = lib().or_regex_find_iter(self.h, regex_index, buf.ctypes.data, len(buf), None, 0)
        out = np.zeros(max(int(n), 1), dtype=MATCH_DTYPE)
        lib().or_regex_find_iter(self.h, regex_index, buf.ctypes.data, len(buf), out.ctypes.data, n)
        return [(int(s), int(e)) for s, e in zip(out["start"][: int(n)], out["end"][: int(n)])]


class StreamScannerOracle:
    """Restatement of StreamScanner (core/src/stream.rs:38-134): each feed
    scans window = carried tail + chunk, keeps the matches that end past the
    tail (stream.rs:104) and shifts them by processed - len(tail)
    (stream.rs:101-108); the new tail is the window's last `keep` bytes
    (stream.rs:110-120), keep = the longest literal pattern (regexes add
    nothing: core/src/pattern/compiler.rs:149,201). finish() is feed(b"")."""

    def __init__(self, oracle: "Oracle", keep: int, mode: int = 1):
        self.oracle = oracle
        self.keep = keep
        self.mode = mode
        self.tail = b""
        self.processed = 0

    def feed(self, chunk):
        chunk = _b(chunk)
        window = self.tail + chunk
        base = self.processed - len(self.tail)
        found = self.oracle.scan_list(window, self.mode) if window else []
        out = [(p, s + base, e + base) for p, s, e in found if e > len(self.tail)]
        keep = min(self.keep, len(window))
        self.tail = window[len(window) - keep:] if keep else b""
        self.processed += len(chunk)
        return out

    def finish(self):
        return self.feed(b"")
