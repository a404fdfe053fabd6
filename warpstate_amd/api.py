"""Python mirror of the reference crate's scan API over the C ABI.

Names and argument meanings follow the reference (paths relative to its
checkout): PatternSet / PatternSetBuilder (core/src/pattern/), GpuMatcher —
documented there as WarpStateer — (core/src/gpu/mod.rs), AutoMatcherConfig
(core/src/config.rs), StreamPipeline (core/src/pipeline.rs) and the batch
free functions (core/src/batch.rs). Every scan runs on the GPU through
libwarpstate_amd.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass, field
from typing import Iterable, List, NamedTuple, Optional, Sequence

import numpy as np

from . import _native as N


class Match(NamedTuple):
    """matchkit::Match: {pattern_id, start, end}, end exclusive."""

    pattern_id: int
    start: int
    end: int


class ScanItem(NamedTuple):
    """batch::ScanItem (core/src/batch.rs:13-18)."""

    id: int
    data: bytes


class TaggedMatch(NamedTuple):
    """batch::TaggedMatch (core/src/batch.rs:22-27)."""

    source_id: int
    matched: Match


def _as_bytes(x) -> bytes:
    if isinstance(x, str):
        return x.encode()
    return bytes(x)


@dataclass
class AutoMatcherConfig:
    """Subset of core/src/config.rs:39-173 that the GPU path honours."""

    max_matches: int = 1_000_000
    chunk_size: int = 128 * 1024 * 1024
    chunk_overlap: int = 4096
    gpu_threshold: int = 64 * 1024

    def with_max_matches(self, n: int) -> "AutoMatcherConfig":
        self.max_matches = n
        return self

    def with_chunk_size(self, n: int) -> "AutoMatcherConfig":
        self.chunk_size = n
        return self

    def with_chunk_overlap(self, n: int) -> "AutoMatcherConfig":
        self.chunk_overlap = n
        return self

    def to_c(self) -> N.WsConfig:
        return N.WsConfig(self.max_matches, self.chunk_size, self.chunk_overlap, 0, 0)


@dataclass
class Tuning:
    """Kernel choice: variant (hash / dfa_global / dfa_lds) and launch shape."""

    variant: int = N.VARIANT_AUTO
    bytes_per_lane: int = 0
    blocks_per_cu: int = 0
    checked: bool = False  # host-checked pipeline instead of the device-sized one

    def to_c(self) -> N.WsTuning:
        return N.WsTuning(self.variant, self.bytes_per_lane, self.blocks_per_cu,
                          N.TUNING_CHECKED if self.checked else 0)


class Context:
    """One device context: a HIP stream plus workspaces (thread-safe, serialised)."""

    def __init__(self, device: int = 0, config: Optional[AutoMatcherConfig] = None):
        lib = N.load()
        self.config = config or AutoMatcherConfig()
        h = ctypes.c_void_p()
        cfg = self.config.to_c()
        N.check(lib.ws_ctx_create(device, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self.device = device

    def close(self) -> None:
        if self.handle:
            N.load().ws_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return N.load().ws_ctx_stream(self.handle) or 0

    @property
    def cu_count(self) -> int:
        return int(N.load().ws_ctx_cu_count(self.handle))

    @property
    def graph_count(self) -> int:
        """Captured launch sequences (hipGraphs) this context replays."""
        return int(N.load().ws_ctx_graph_count(self.handle))

    def set_profiling(self, on: bool) -> None:
        N.load().ws_ctx_set_profiling(self.handle, 1 if on else 0)

    def set_graphs(self, on: bool) -> None:
        """Capture and replay repeated device-sized calls (default on)."""
        N.load().ws_ctx_set_graphs(self.handle, 1 if on else 0)

    def last_timings(self) -> N.WsTimings:
        t = N.WsTimings()
        N.check(N.load().ws_ctx_last_timings(self.handle, ctypes.byref(t)))
        return t


_default_ctx = {}
_default_lock = threading.Lock()


def default_context(device: int = 0) -> Context:
    with _default_lock:
        if device not in _default_ctx:
            _default_ctx[device] = Context(device)
        return _default_ctx[device]


class PatternSetBuilder:
    """PatternSetBuilder (core/src/pattern/compiler.rs:13-93)."""

    def __init__(self):
        self._patterns: List[tuple] = []
        self._ci = False

    def literal(self, s) -> "PatternSetBuilder":
        self._patterns.append((0, _as_bytes(s), None))
        return self

    def literal_bytes(self, b) -> "PatternSetBuilder":
        return self.literal(b)

    def named_literal(self, name: str, s) -> "PatternSetBuilder":
        self._patterns.append((0, _as_bytes(s), name))
        return self

    def regex(self, s) -> "PatternSetBuilder":
        self._patterns.append((1, _as_bytes(s), None))
        return self

    def named_regex(self, name: str, s) -> "PatternSetBuilder":
        self._patterns.append((1, _as_bytes(s), name))
        return self

    def case_insensitive(self, yes: bool = True) -> "PatternSetBuilder":
        self._ci = yes
        return self

    def build(self) -> "PatternSet":
        ps = PatternSet(list(self._patterns), self._ci)
        ps.analyze()  # host-side compile: errors surface at build time, no GPU needed
        return ps


class PatternSet:
    """A compiled pattern set. `scan` has PatternSet::scan semantics
    (core/src/pattern/mod.rs:112-140) and `scan_overlapping` has
    PatternSet::scan_overlapping semantics (mod.rs:189-219); both run on the GPU."""

    def __init__(self, patterns: List[tuple], case_insensitive: bool = False):
        self.patterns = patterns
        self.case_insensitive = case_insensitive
        self._handles = {}
        self._lock = threading.Lock()
        self._index: Optional[bytes] = None  # set for a set loaded from a compiled index
        self._index_info: Optional[dict] = None

    @classmethod
    def _from_index(cls, data: bytes, header: dict, info: dict) -> "PatternSet":
        ps = cls([], bool(header["case_insensitive"]))
        ps._index = data
        ps._index_info = info
        return ps

    @staticmethod
    def builder() -> PatternSetBuilder:
        return PatternSetBuilder()

    def __len__(self) -> int:
        if self._index is not None:
            return int(self._index_info["pattern_count"])
        return len(self.patterns)

    def _compiled(self, ctx: Context):
        # A compiled set is read-only device data: every context on the same
        # device shares one copy.
        key = ctx.device
        with self._lock:
            h = self._handles.get(key)
            if h is None:
                lib = N.load()
                h = ctypes.c_void_p()
                if self._index is not None:
                    N.check(lib.ws_index_load(ctx.handle, self._index, len(self._index), ctypes.byref(h)))
                else:
                    arr, _bufs = self._c_patterns()
                    flags = 1 if self.case_insensitive else 0
                    N.check(lib.ws_patterns_compile(ctx.handle, arr, len(self.patterns), flags, ctypes.byref(h)))
                self._handles[key] = (h, ctx)
                return h
            return h[0]

    def _c_patterns(self):
        bufs = [ctypes.create_string_buffer(b, len(b)) if b else None for _, b, _ in self.patterns]
        arr = (N.WsPattern * len(self.patterns))()
        for i, ((kind, b, name), buf) in enumerate(zip(self.patterns, bufs)):
            arr[i].kind = kind
            arr[i].bytes = ctypes.cast(buf, ctypes.c_void_p) if buf is not None else None
            arr[i].len = len(b)
            arr[i].name = name.encode() if name else None
        return arr, bufs

    def analyze(self) -> dict:
        """Host-only compile (ws_patterns_analyze): validation + set facts."""
        if self._index is not None:
            return dict(self._index_info)
        arr, _bufs = self._c_patterns()
        i = N.WsPatternsInfo()
        N.check(N.load().ws_patterns_analyze(arr, len(self.patterns), 1 if self.case_insensitive else 0,
                                             ctypes.byref(i)))
        return {name: getattr(i, name) for name, _ in N.WsPatternsInfo._fields_}

    def info(self, ctx: Optional[Context] = None) -> dict:
        h = self._compiled(ctx or default_context())
        i = N.WsPatternsInfo()
        N.check(N.load().ws_patterns_info_get(h, ctypes.byref(i)))
        return {name: getattr(i, name) for name, _ in N.WsPatternsInfo._fields_}

    def max_pattern_len(self) -> int:
        return int(self.analyze()["max_pattern_len"])

    def names(self, ctx: Optional[Context] = None) -> List[Optional[str]]:
        """Per pattern id, its name or None (from the compiled set)."""
        lib = N.load()
        h = self._compiled(ctx or default_context())
        out: List[Optional[str]] = []
        for i in range(len(self)):
            n = ctypes.c_uint64(0)
            N.check(lib.ws_patterns_name(h, i, None, 0, ctypes.byref(n)))
            if n.value == 2**64 - 1:
                out.append(None)
                continue
            buf = ctypes.create_string_buffer(max(1, n.value))
            N.check(lib.ws_patterns_name(h, i, buf, n.value, ctypes.byref(n)))
            out.append(buf.raw[: n.value].decode())
        return out

    def scan(self, data) -> List[Match]:
        return GpuMatcher(self).scan(data, mode="standard")

    def scan_overlapping(self, data) -> List[Match]:
        return GpuMatcher(self).scan(data, mode="overlapping")

    def scan_count(self, data) -> int:
        return len(self.scan(data))

    def __del__(self):
        try:
            lib = N.load()
            for h, _ctx in self._handles.values():
                lib.ws_patterns_free(h)
        except Exception:
            pass


class CompiledPatternIndex:
    """CompiledPatternIndex (core/src/compiled_index/mod.rs:22-845): a
    serialized pattern set (WPSIDX01). `build` writes version 5 by default
    (the reference's sections plus the compiled DFA and automaton tables, so
    loading skips compilation) or version 3 (`version=3`, sets without
    regexes: a file the reference crate loads). `load` reads versions 1-5;
    files of versions 1-4 are rebuilt from their stored patterns as
    to_pattern_set does. `scan` runs to_pattern_set().scan on the GPU: the
    reference's index scanner returns PatternSet::scan's results (its
    round-trip tests, core/src/compiled_index/tests.rs:5-128)."""

    MAX_FILE_BYTES = 512 * 1024 * 1024  # mod.rs:19

    def __init__(self, data: bytes, header: dict, info: dict):
        self._data = data
        self._header = header
        self._info = info
        self._set: Optional[PatternSet] = None

    @staticmethod
    def build(patterns: "PatternSet", version: int = N.INDEX_VERSION) -> bytes:
        """CompiledPatternIndex::build (builder.rs:8-61); host only."""
        flags = N.INDEX_REFERENCE_V3 if version == 3 else 0
        if version not in (3, N.INDEX_VERSION):
            raise ValueError(f"index version {version} cannot be written (3 or {N.INDEX_VERSION})")
        lib = N.load()
        n = ctypes.c_uint64(0)
        if patterns._index is not None:
            h = patterns._compiled(default_context())
            N.check(lib.ws_patterns_serialize(h, flags, None, 0, ctypes.byref(n)))
            out = ctypes.create_string_buffer(n.value)
            N.check(lib.ws_patterns_serialize(h, flags, out, n.value, ctypes.byref(n)))
            return out.raw[: n.value]
        arr, _bufs = patterns._c_patterns()
        ci = 1 if patterns.case_insensitive else 0
        # One compilation when the first guess at the size suffices.
        cap = 1 << 16
        for p in patterns.patterns:
            cap += 64 + 32 * len(p[1])
        while True:
            out = ctypes.create_string_buffer(cap)
            st = lib.ws_index_build(arr, len(patterns.patterns), ci, flags, out, cap, ctypes.byref(n))
            if st == 13 and n.value > cap:
                cap = int(n.value)
                continue
            N.check(st)
            return out.raw[: n.value]

    @classmethod
    def load(cls, data) -> "CompiledPatternIndex":
        """CompiledPatternIndex::load (mod.rs:58-61): checks the whole file on the host."""
        data = bytes(data)
        hdr = N.WsIndexInfo()
        info = N.WsPatternsInfo()
        N.check(N.load().ws_index_inspect(data if data else None, len(data), ctypes.byref(hdr), ctypes.byref(info)))
        return cls(data, {k: getattr(hdr, k) for k, _ in N.WsIndexInfo._fields_},
                   {k: getattr(info, k) for k, _ in N.WsPatternsInfo._fields_})

    @classmethod
    def load_from_file(cls, path) -> "CompiledPatternIndex":
        """CompiledPatternIndex::load_from_file (mod.rs:76-94)."""
        import os

        size = os.stat(path).st_size
        if size > cls.MAX_FILE_BYTES:
            raise N.WarpstateError(10, f"compiled index is {size} bytes; the limit is {cls.MAX_FILE_BYTES}")
        with open(path, "rb") as f:
            return cls.load(f.read())

    def save_to_file(self, path) -> None:
        """CompiledPatternIndex::save_to_file (mod.rs:65-73)."""
        N.check(N.load().ws_index_save_file(self._data, len(self._data), str(path).encode()))

    def to_bytes(self) -> bytes:
        return self._data

    @property
    def version(self) -> int:
        return int(self._header["version"])

    @property
    def prebuilt(self) -> bool:
        return bool(self._header["prebuilt"])

    def literal_count(self) -> int:
        return int(self._header["literal_count"])

    def to_pattern_set(self) -> "PatternSet":
        """CompiledPatternIndex::to_pattern_set (mod.rs:287-372); compiled on
        the device from the file's tables when the file carries them."""
        if self._set is None:
            self._set = PatternSet._from_index(self._data, self._header, self._info)
        return self._set

    def names(self) -> List[Optional[str]]:
        return self.to_pattern_set().names()

    def scan(self, data) -> List[Match]:
        return self.to_pattern_set().scan(data)


def _mode_code(mode: str) -> int:
    if mode in ("overlapping", N.MODE_OVERLAPPING):
        return N.MODE_OVERLAPPING
    if mode in ("standard", N.MODE_STANDARD):
        return N.MODE_STANDARD
    raise ValueError(f"unknown mode {mode!r}")


def matches_to_list(arr: np.ndarray) -> List[Match]:
    return [Match(int(a), int(b), int(c)) for a, b, c in zip(arr["pattern_id"], arr["start"], arr["end"])]


class GpuMatcher:
    """GpuMatcher (core/src/gpu/mod.rs:95-306); the reference README calls it
    WarpStateer. `scan` returns the GPU contract: overlapping literal matches
    plus leftmost-first regex matches, sorted by (start, pattern_id, end)."""

    def __init__(self, patterns: PatternSet, config: Optional[AutoMatcherConfig] = None,
                 context: Optional[Context] = None, device: int = 0):
        if context is None:
            context = default_context(device) if config is None else Context(device, config)
        self.ctx = context
        self.patterns = patterns
        self.config = context.config
        self.handle = patterns._compiled(context)

    def max_block_size(self) -> int:
        return self.config.chunk_size

    def scan_array(self, data, mode="overlapping", tuning: Optional[Tuning] = None,
                   capacity: Optional[int] = None) -> np.ndarray:
        """Host input → structured numpy array (MATCH_DTYPE)."""
        lib = N.load()
        buf = np.frombuffer(_as_bytes(data) if not isinstance(data, np.ndarray) else data.tobytes(), dtype=np.uint8)
        cap = capacity if capacity is not None else max(1024, min(self.config.max_matches, len(buf) // 8 + 1024))
        t = tuning.to_c() if tuning else None
        while True:
            out = np.zeros(cap, dtype=N.MATCH_DTYPE)
            count = ctypes.c_uint64(0)
            st = lib.ws_scan(self.ctx.handle, self.handle, buf.ctypes.data if len(buf) else None, len(buf),
                             N.MEM_HOST, _mode_code(mode), ctypes.byref(t) if t else None,
                             out.ctypes.data, cap, N.MEM_HOST, ctypes.byref(count))
            if st == 11 and capacity is None and count.value > cap and count.value <= self.config.max_matches:
                cap = int(count.value)
                continue
            N.check(st, int(count.value))
            return out[: count.value]

    def scan(self, data, mode="overlapping", tuning: Optional[Tuning] = None) -> List[Match]:
        return matches_to_list(self.scan_array(data, mode, tuning))

    def scan_device(self, data_ptr: int, length: int, out_ptr: int, capacity: int, mode="overlapping",
                    tuning: Optional[Tuning] = None, out_on_device: bool = True) -> int:
        """Device-resident input (and output) — the HBM-resident fast path."""
        lib = N.load()
        count = ctypes.c_uint64(0)
        t = tuning.to_c() if tuning else None
        st = lib.ws_scan(self.ctx.handle, self.handle, ctypes.c_void_p(data_ptr), length, N.MEM_DEVICE,
                         _mode_code(mode), ctypes.byref(t) if t else None, ctypes.c_void_p(out_ptr), capacity,
                         N.MEM_DEVICE if out_on_device else N.MEM_HOST, ctypes.byref(count))
        N.check(st, int(count.value))
        return int(count.value)

    def scan_device_async(self, data_ptr: int, length: int, out_ptr: int, capacity: int, mode="overlapping",
                          tuning: Optional[Tuning] = None, out_on_device: bool = True) -> "PendingScan":
        """scan_device without waiting: the scan is enqueued and a PendingScan
        returned (GpuMatcher::scan's future, core/src/gpu/mod.rs:223-260).
        The buffers must stay valid until wait(); one outstanding call per
        context."""
        t = tuning.to_c() if tuning else None
        st = N.load().ws_scan_async(self.ctx.handle, self.handle, ctypes.c_void_p(data_ptr), length, N.MEM_DEVICE,
                                    _mode_code(mode), ctypes.byref(t) if t else None, ctypes.c_void_p(out_ptr),
                                    capacity, N.MEM_DEVICE if out_on_device else N.MEM_HOST)
        N.check(st)
        return PendingScan(self)

    def prepare_device(self, data_ptr: int, length: int, out_ptr: int, capacity: int, mode="overlapping",
                       tuning: Optional[Tuning] = None) -> "PreparedScan":
        """scan_device with its arguments converted once: a repeated call on
        the same buffers costs one foreign call (about 10 us less per call
        from Python than scan_device)."""
        return PreparedScan(self, data_ptr, length, out_ptr, capacity, mode, tuning)

    def scan_device_array(self, data_ptr: int, length: int, mode="overlapping", tuning: Optional[Tuning] = None,
                          capacity: Optional[int] = None) -> np.ndarray:
        """Device-resident input, match list returned to the host (MATCH_DTYPE)."""
        lib = N.load()
        cap = capacity if capacity is not None else max(1024, min(self.config.max_matches, length // 64 + 4096))
        t = tuning.to_c() if tuning else None
        while True:
            out = np.empty(cap, dtype=N.MATCH_DTYPE)
            count = ctypes.c_uint64(0)
            st = lib.ws_scan(self.ctx.handle, self.handle, ctypes.c_void_p(data_ptr), length, N.MEM_DEVICE,
                             _mode_code(mode), ctypes.byref(t) if t else None, out.ctypes.data, cap, N.MEM_HOST,
                             ctypes.byref(count))
            if st == 11 and capacity is None and count.value > cap and count.value <= self.config.max_matches:
                cap = int(count.value)
                continue
            N.check(st, int(count.value))
            return out[: count.value]

    def info(self) -> dict:
        return self.patterns.info(self.ctx)


WarpStateer = GpuMatcher


class PreparedScan:
    """A device-resident scan (ws_scan on device input and output) whose
    foreign-call arguments are built once; calling it returns the match count
    and raises as scan_device does."""

    def __init__(self, matcher: GpuMatcher, data_ptr: int, length: int, out_ptr: int, capacity: int,
                 mode="overlapping", tuning: Optional[Tuning] = None):
        self._fn = N.load().ws_scan
        self._count = ctypes.c_uint64(0)
        self._tuning = tuning.to_c() if tuning else None
        self._args = (matcher.ctx.handle, matcher.handle, ctypes.c_void_p(data_ptr), ctypes.c_uint64(length),
                      ctypes.c_uint32(N.MEM_DEVICE), ctypes.c_uint32(_mode_code(mode)),
                      ctypes.byref(self._tuning) if self._tuning is not None else None, ctypes.c_void_p(out_ptr),
                      ctypes.c_uint64(capacity), ctypes.c_uint32(N.MEM_DEVICE), ctypes.byref(self._count))
        self._keep = matcher  # the compiled set and context outlive the prepared call

    def __call__(self) -> int:
        st = self._fn(*self._args)
        if st:
            N.check(st, int(self._count.value))
        return self._count.value


class PendingScan:
    """An outstanding GpuMatcher.scan_device_async call: ready() polls,
    wait() returns the match count (raising as scan_device would)."""

    def __init__(self, matcher: GpuMatcher):
        self._m = matcher  # keeps the context and the compiled set alive
        self._done = False
        self._count = 0
        self._error = None

    def ready(self) -> bool:
        return self._done or N.load().ws_poll(self._m.ctx.handle) == 1

    def wait(self) -> int:
        if not self._done:
            count = ctypes.c_uint64(0)
            st = N.load().ws_wait(self._m.ctx.handle, ctypes.byref(count))
            self._done = True
            self._count = int(count.value)
            try:
                N.check(st, self._count)
            except N.WarpstateError as e:
                self._error = e
        if self._error is not None:
            raise self._error
        return self._count


class StreamPipeline:
    """StreamPipeline (core/src/pipeline.rs:16-153): chunked scanning with an
    overlap window. Unlike the reference's dedup stitching, regex chains carry
    across chunks, so the result equals a whole-input scan."""

    def __init__(self, backend: GpuMatcher, overlap_window: int):
        self.backend = backend
        self.overlap_window = overlap_window

    def scan_array(self, data, mode="overlapping", tuning: Optional[Tuning] = None,
                   chunk_size: Optional[int] = None) -> np.ndarray:
        lib = N.load()
        buf = np.frombuffer(_as_bytes(data), dtype=np.uint8)
        chunk = chunk_size or self.backend.max_block_size()
        cap = max(1024, min(self.backend.config.max_matches, len(buf) // 8 + 1024))
        t = tuning.to_c() if tuning else None
        while True:
            out = np.zeros(cap, dtype=N.MATCH_DTYPE)
            count = ctypes.c_uint64(0)
            st = lib.ws_stream_scan(self.backend.ctx.handle, self.backend.handle,
                                    buf.ctypes.data if len(buf) else None, len(buf), N.MEM_HOST, chunk,
                                    self.overlap_window, _mode_code(mode), ctypes.byref(t) if t else None,
                                    out.ctypes.data, cap, N.MEM_HOST, ctypes.byref(count))
            if st == 11 and count.value > cap and count.value <= self.backend.config.max_matches:
                cap = int(count.value)
                continue
            N.check(st, int(count.value))
            return out[: count.value]

    def scan_device(self, data_ptr: int, length: int, out_ptr: int, capacity: int, mode="overlapping",
                    tuning: Optional[Tuning] = None, chunk_size: Optional[int] = None,
                    out_on_device: bool = True) -> int:
        """Device-resident input; the match list goes to out_ptr (device or host)."""
        lib = N.load()
        chunk = chunk_size or self.backend.max_block_size()
        count = ctypes.c_uint64(0)
        t = tuning.to_c() if tuning else None
        st = lib.ws_stream_scan(self.backend.ctx.handle, self.backend.handle, ctypes.c_void_p(data_ptr), length,
                                N.MEM_DEVICE, chunk, self.overlap_window, _mode_code(mode),
                                ctypes.byref(t) if t else None, ctypes.c_void_p(out_ptr), capacity,
                                N.MEM_DEVICE if out_on_device else N.MEM_HOST, ctypes.byref(count))
        N.check(st, int(count.value))
        return int(count.value)

    def scan(self, data, chunk_size: Optional[int] = None) -> List[Match]:
        return matches_to_list(self.scan_array(data, chunk_size=chunk_size))


class HotSwapPatternSet:
    """HotSwapPatternSet (core/src/pattern/mod.rs:17-70): the active compiled
    set, replaced atomically by swap(); a scan keeps the set that was active
    when it started. Thread-safe: scans from several threads (each with its
    own Context, or sharing one) may run while another thread swaps."""

    def __init__(self, patterns: "PatternSet", device: int = 0, config: Optional[AutoMatcherConfig] = None):
        self.config = config or AutoMatcherConfig()
        self.ctx = Context(device, self.config)
        h = ctypes.c_void_p()
        N.check(N.load().ws_hotswap_create(patterns._compiled(self.ctx), ctypes.byref(h)))
        self.handle = h.value

    def swap(self, nxt: "PatternSet") -> bool:
        """True if `nxt` replaced the active set, False if it already was it.
        The hot-swap slot holds its own reference to the compiled set."""
        return bool(N.load().ws_hotswap_swap(self.handle, nxt._compiled(self.ctx)))

    def scan_array(self, data, mode="standard", ctx: Optional[Context] = None) -> np.ndarray:
        lib = N.load()
        c = ctx or self.ctx
        buf = np.frombuffer(_as_bytes(data), dtype=np.uint8)
        cap = max(1024, min(self.config.max_matches, len(buf) // 8 + 1024))
        while True:
            out = np.zeros(cap, dtype=N.MATCH_DTYPE)
            count = ctypes.c_uint64(0)
            st = lib.ws_hotswap_scan(c.handle, self.handle, buf.ctypes.data if len(buf) else None, len(buf),
                                     N.MEM_HOST, _mode_code(mode), None, out.ctypes.data, cap, N.MEM_HOST,
                                     ctypes.byref(count))
            if st == 11 and count.value > cap and count.value <= self.config.max_matches:
                cap = int(count.value)
                continue
            N.check(st, int(count.value))
            return out[: count.value]

    def scan(self, data, mode="standard", ctx: Optional[Context] = None) -> List[Match]:
        """PatternSet::scan semantics by default, as HotSwapPatternSet::scan."""
        return matches_to_list(self.scan_array(data, mode, ctx))

    def close(self) -> None:
        if getattr(self, "handle", None):
            N.load().ws_hotswap_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StreamScanner:
    """StreamScanner (core/src/stream.rs:14-135): incremental feeds. Each feed
    scans the carried tail plus the new chunk on the device and returns the
    matches that end past the tail, at stream offsets; the tail is the last
    L bytes, L the longest literal pattern. mode "standard" (PatternSet::scan
    semantics) is the reference's; "overlapping" reports every occurrence."""

    def __init__(self, backend, mode: str = "standard"):
        matcher = backend if isinstance(backend, GpuMatcher) else GpuMatcher(backend)
        self.backend = matcher
        lib = N.load()
        h = ctypes.c_void_p()
        N.check(lib.ws_stream_scanner_create(matcher.ctx.handle, matcher.handle, _mode_code(mode), ctypes.byref(h)))
        self.handle = h.value

    def _run(self, fn, *args) -> List[Match]:
        lib = N.load()
        cap = 1024
        while True:
            out = np.zeros(cap, dtype=N.MATCH_DTYPE)
            count = ctypes.c_uint64(0)
            st = fn(self.handle, *args, out.ctypes.data, cap, ctypes.byref(count))
            if st == 11 and count.value > cap and count.value <= self.backend.config.max_matches:
                cap = int(count.value)
                continue
            N.check(st, int(count.value))
            return matches_to_list(out[: count.value])

    def feed(self, chunk) -> List[Match]:
        lib = N.load()
        buf = np.frombuffer(_as_bytes(chunk), dtype=np.uint8)
        return self._run(lib.ws_stream_scanner_feed, buf.ctypes.data if len(buf) else None, len(buf), N.MEM_HOST)

    def feed_device(self, data_ptr: int, length: int) -> List[Match]:
        """A device-resident chunk."""
        lib = N.load()
        return self._run(lib.ws_stream_scanner_feed, ctypes.c_void_p(data_ptr), length, N.MEM_DEVICE)

    def finish(self) -> List[Match]:
        return self._run(N.load().ws_stream_scanner_finish)

    @property
    def processed_bytes(self) -> int:
        return int(N.load().ws_stream_scanner_processed(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            N.load().ws_stream_scanner_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _batch_array(matcher: GpuMatcher, items: Sequence[ScanItem], mode: str,
                 tuning: Optional[Tuning] = None) -> np.ndarray:
    lib = N.load()
    keep = []
    arr = (N.WsItem * len(items))()
    total = 0
    for i, it in enumerate(items):
        b = np.frombuffer(_as_bytes(it.data), dtype=np.uint8)
        keep.append(b)
        arr[i].id = int(it.id)
        arr[i].data = b.ctypes.data if len(b) else None
        arr[i].len = len(b)
        total += len(b)
    cap = max(1024, min(matcher.config.max_matches, total // 8 + 1024))
    t = tuning.to_c() if tuning else None
    while True:
        out = np.zeros(cap, dtype=N.TAGGED_DTYPE)
        count = ctypes.c_uint64(0)
        st = lib.ws_scan_batch(matcher.ctx.handle, matcher.handle, arr, len(items), N.MEM_HOST,
                               _mode_code(mode), ctypes.byref(t) if t else None, out.ctypes.data, cap,
                               ctypes.byref(count))
        if st == 11 and count.value > cap and count.value <= matcher.config.max_matches:
            cap = int(count.value)
            continue
        N.check(st, int(count.value))
        return out[: count.value]


def scan_batch_device(matcher: GpuMatcher, ids: np.ndarray, ptrs: np.ndarray, lens: np.ndarray,
                      mode: str = "overlapping", tuning: Optional[Tuning] = None,
                      capacity: Optional[int] = None) -> np.ndarray:
    """scan_batch over items already in device memory (one allocation span of
    at most 4 GiB): ids u64, ptrs device addresses, lens bytes. Returns
    TAGGED_DTYPE records in (source_id, start, pattern_id, end) order."""
    lib = N.load()
    n = len(ids)
    arr = (N.WsItem * n)()
    view = np.frombuffer(arr, dtype=np.dtype([("id", "<u8"), ("data", "<u8"), ("len", "<u8")]), count=n)
    view["id"] = ids
    view["data"] = ptrs
    view["len"] = lens
    total = int(np.sum(lens))
    # Output capacity: a quarter above the matcher's previous batch count
    # once there is one (a bound of one record per 64 input bytes otherwise).
    # A modest buffer comes from the allocator's reused heap pages, where a
    # bound-sized one would be fresh pages that the copy faults in.
    hint = getattr(matcher, "_batch_count_hint", None)
    bound = max(1024, min(matcher.config.max_matches, total // 64 + 4096))
    if capacity is not None:
        cap = capacity
    elif hint is not None:
        cap = max(1024, min(bound, hint + hint // 4 + 4096))
    else:
        cap = bound
    t = tuning.to_c() if tuning else None
    while True:
        out = np.empty(cap, dtype=N.TAGGED_DTYPE)
        count = ctypes.c_uint64(0)
        st = lib.ws_scan_batch(matcher.ctx.handle, matcher.handle, arr, n, N.MEM_DEVICE, _mode_code(mode),
                               ctypes.byref(t) if t else None, out.ctypes.data, cap, ctypes.byref(count))
        if st == 11 and capacity is None and count.value > cap and count.value <= matcher.config.max_matches:
            cap = int(count.value)
            continue
        N.check(st, int(count.value))
        if capacity is None:
            matcher._batch_count_hint = int(count.value)
        return out[: count.value]


class BatchPlan:
    """A batch of device-resident items laid out once (batch::coalesce and
    planned_batches, core/src/batch.rs:65-104,197-236): id order, segments,
    work units and dense positions live on the device, so each
    `scan_batch_planned` call over it does no host planning, is device-sized
    and, when repeated with the same output, replays a captured hipGraph.
    The items must stay where they are while the plan is used."""

    def __init__(self, ctx: Context, ids: np.ndarray, ptrs: np.ndarray, lens: np.ndarray):
        n = len(ids)
        arr = (N.WsItem * n)()
        view = np.frombuffer(arr, dtype=np.dtype([("id", "<u8"), ("data", "<u8"), ("len", "<u8")]), count=n)
        view["id"] = ids
        view["data"] = ptrs
        view["len"] = lens
        h = ctypes.c_void_p()
        N.check(N.load().ws_batch_plan_create(ctx.handle, arr, n, ctypes.byref(h)))
        self.h = h
        self.ctx = ctx
        self.n_items = n
        self.total = int(np.sum(lens))

    def close(self) -> None:
        if getattr(self, "h", None) and self.h.value:
            N.load().ws_batch_plan_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def scan_batch_planned(matcher: GpuMatcher, plan: BatchPlan, mode: str = "overlapping",
                       tuning: Optional[Tuning] = None, out_ptr: Optional[int] = None,
                       capacity: Optional[int] = None):
    """scan_batch over a BatchPlan. With out_ptr (a device address holding
    `capacity` TAGGED_DTYPE records) the matches stay on the device and the
    count is returned; otherwise a TAGGED_DTYPE array comes back."""
    lib = N.load()
    t = tuning.to_c() if tuning else None
    count = ctypes.c_uint64(0)
    if out_ptr is not None:
        st = lib.ws_scan_batch_planned(matcher.ctx.handle, matcher.handle, plan.h, _mode_code(mode),
                                       ctypes.byref(t) if t else None, ctypes.c_void_p(out_ptr), capacity,
                                       N.MEM_DEVICE, ctypes.byref(count))
        N.check(st, int(count.value))
        return int(count.value)
    cap = capacity if capacity is not None else max(1024, min(matcher.config.max_matches, plan.total // 64 + 4096))
    while True:
        out = np.empty(cap, dtype=N.TAGGED_DTYPE)
        st = lib.ws_scan_batch_planned(matcher.ctx.handle, matcher.handle, plan.h, _mode_code(mode),
                                       ctypes.byref(t) if t else None, out.ctypes.data, cap, N.MEM_HOST,
                                       ctypes.byref(count))
        if st == 11 and capacity is None and count.value > cap and count.value <= matcher.config.max_matches:
            cap = int(count.value)
            continue
        N.check(st, int(count.value))
        return out[: count.value]


def _tagged_list(arr: np.ndarray) -> List[TaggedMatch]:
    return [TaggedMatch(int(s), Match(int(p), int(a), int(b)))
            for s, p, a, b in zip(arr["source_id"], arr["pattern_id"], arr["start"], arr["end"])]


def scan_batch_gpu(matcher: GpuMatcher, items: Iterable[ScanItem]) -> List[TaggedMatch]:
    """batch::scan_batch_gpu (core/src/batch.rs:336-363), per-item exact:
    overlapping literals + leftmost-first regexes, no cross-item matches."""
    return _tagged_list(_batch_array(matcher, list(items), "overlapping"))


def scan_batch(matcher: GpuMatcher, items: Iterable[ScanItem]) -> List[TaggedMatch]:
    """batch::scan_batch (core/src/batch.rs:314-320)."""
    return scan_batch_gpu(matcher, items)


def coalesce(items: Sequence[ScanItem]):
    """batch::coalesce (core/src/batch.rs:65-98): (entries, buffer)."""
    entries, buf, off = [], bytearray(), 0
    for it in items:
        d = _as_bytes(it.data)
        entries.append((it.id, off, len(d)))
        buf += d
        off += len(d)
    return entries, bytes(buf)
