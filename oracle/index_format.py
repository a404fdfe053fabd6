"""CPU restatement of the reference's compiled-index format (WPSIDX01) —
TEST INFRASTRUCTURE ONLY.

Used by tests/ as the checker of the library's index reader and writer; the
product package never imports it.

* `crc32_ieee`   — core/src/compiled_index/builder.rs:65-78 (bitwise, reflected
                   0xEDB88320, initial and final inversion).
* `build`        — CompiledPatternIndex::build, builder.rs:8-61 and the push_*
                   helpers builder.rs:80-180, for versions 1, 2 and 3. The
                   reference's regex sections hold regex-automata DFAs; this
                   restatement writes sections of the same shape with
                   placeholder contents (empty tables, no native bytes), which
                   is what the reader sees of them.
* `parse`        — CompiledPatternIndex::parse, mod.rs:96-271, and the Cursor
                   reads, mod.rs:466-765; raises IndexFormatError with the
                   reference's wording for each check.
* `to_pattern_list` — CompiledPatternIndex::to_pattern_set, mod.rs:287-372:
                   the pattern list the reference rebuilds from a file.

Patterns are given as resolved (kind, bytes, name) triples: kind 0 for
what the reference stores as a literal (literals and regexes that are plain
literals, compiler.rs:172-211), kind 1 for a regex kept as a regex.
"""
from __future__ import annotations

import struct
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .oracle import fnv1a_window, prefilter_table

MAGIC = b"WPSIDX01"   # mod.rs:17
REF_VERSION = 3       # mod.rs:18
OWN_VERSION = 5       # the library's version
OWN_VERSIONS = (4, 5)  # the library's layouts (4: an earlier one of its regex sections)
HASH_WINDOW = 8       # compiler.rs:118
NO_NAME = 0xFFFFFFFF  # builder.rs:128-140


class IndexFormatError(Exception):
    pass


def crc32_ieee(data: bytes) -> int:
    crc = 0xFFFFFFFF
    for b in data:
        crc ^= b
        for _ in range(8):
            crc = (crc >> 1) ^ 0xEDB88320 if crc & 1 else crc >> 1
    return crc ^ 0xFFFFFFFF


def _u32(v: int) -> bytes:
    return struct.pack("<I", v)


def build(patterns: Sequence[Tuple[int, bytes, Optional[str]]], case_insensitive: bool = False,
          version: int = REF_VERSION) -> bytes:
    lits = [(i, p) for i, (k, p, _) in enumerate(patterns) if k == 0]
    rxs = [(i, p) for i, (k, p, _) in enumerate(patterns) if k == 1]
    out = bytearray(MAGIC)
    out += _u32(version) + _u32(1 if case_insensitive else 0) + _u32(len(patterns))
    out += _u32(len(lits)) + _u32(len(rxs)) + _u32(HASH_WINDOW)
    packed = b"".join(p for _, p in lits)
    out += _u32(len(packed)) + packed                       # push_bytes
    out += _u32(len(lits))                                  # push_u32_pairs
    at = 0
    for _, p in lits:
        out += _u32(at) + _u32(len(p))
        at += len(p)
    out += _u32(len(patterns))                              # push_names
    for _, _, name in patterns:
        if name is None:
            out += _u32(NO_NAME)
        else:
            raw = name.encode()
            out += _u32(len(raw)) + raw
    if version >= 2:                                        # push_regex_patterns
        out += _u32(len(rxs))
        for pid, src in rxs:
            out += _u32(pid) + _u32(len(src)) + src
    out += _u32(len(lits)) + b"".join(_u32(pid) for pid, _ in lits)  # push_usize_slice
    # push_prefilter_table over LiteralPrefilterTable::build
    lens = [len(p) for _, p in lits]
    hashes = [fnv1a_window(p, HASH_WINDOW) for _, p in lits]
    if lits:
        meta, ranges, entries = prefilter_table(lens, hashes, HASH_WINDOW)
    else:
        meta = np.zeros(4 * HASH_WINDOW, np.uint32)
        ranges = entries = np.zeros(0, np.uint32)
    out += _u32(HASH_WINDOW) + meta.astype("<u4").tobytes()
    out += _u32(len(ranges) // 2) + ranges.astype("<u4").tobytes()
    out += _u32(len(entries) // 2) + entries.astype("<u4").tobytes()
    for _ in rxs:                                           # one DFA section per regex
        out += _u32(0) * 4                                  # four empty u32 vectors
        out += _u32(0) * 3                                  # start_state, class_count, eoi_class
        out += _u32(0) * 256                                # byte_classes
        out += _u32(0)                                      # native DFA bytes (none)
        out += _u32(0)                                      # native_original_ids
    if version == REF_VERSION:
        out += _u32(crc32_ieee(bytes(out)))
    return bytes(out)


class _Cursor:
    def __init__(self, data: bytes):
        self.d = data
        self.off = 0

    def remaining(self) -> int:
        return len(self.d) - self.off

    def take(self, n: int) -> bytes:
        if n > self.remaining():
            raise IndexFormatError("compiled index is truncated")
        b = self.d[self.off:self.off + n]
        self.off += n
        return b

    def u32(self) -> int:
        return struct.unpack("<I", self.take(4))[0]

    def fits(self, count: int, size: int, what: str):
        if count * size > self.remaining():
            raise IndexFormatError(f"compiled index {what} declares {count * size} bytes but only "
                                   f"{self.remaining()} bytes remain")

    def vec(self, what: str = "u32 vector") -> List[int]:
        n = self.u32()
        self.fits(n, 4, what)
        return list(struct.unpack(f"<{n}I", self.take(4 * n)))

    def lp_bytes(self) -> bytes:
        return self.take(self.u32())


def parse(data: bytes) -> dict:
    data = bytes(data)
    end = len(data)
    if len(data) >= 16:
        v = struct.unpack_from("<I", data, 8)[0]
        if v == REF_VERSION or v in OWN_VERSIONS:
            stored = struct.unpack_from("<I", data, len(data) - 4)[0]
            if stored != crc32_ieee(data[:-4]):
                raise IndexFormatError("compiled index integrity check failed (CRC mismatch)")
            end = len(data) - 4
    c = _Cursor(data[:end])
    if c.take(8) != MAGIC:
        raise IndexFormatError("compiled index header is invalid")
    version = c.u32()
    if version > OWN_VERSION:
        raise IndexFormatError(f"compiled index version {version} is unsupported")
    if version == 0:
        raise IndexFormatError("compiled index version is invalid")
    ci = c.u32() != 0
    npat, nlit, nrx, window = c.u32(), c.u32(), c.u32(), c.u32()
    packed = c.lp_bytes()
    noff = c.u32()
    c.fits(noff, 8, "pair vector")
    offsets = [(c.u32(), c.u32()) for _ in range(noff)]
    if noff != nlit:
        raise IndexFormatError("compiled index literal metadata is inconsistent")
    nn = c.u32()
    c.fits(nn, 4, "name table")
    names: List[Optional[str]] = []
    for _ in range(nn):
        ln = c.u32()
        if ln == NO_NAME:
            names.append(None)
            continue
        try:
            names.append(c.take(ln).decode("utf-8"))
        except UnicodeDecodeError:
            raise IndexFormatError("compiled index name is not valid UTF-8")
    if nn != npat:
        raise IndexFormatError("compiled index name metadata is inconsistent")
    regexes: Optional[List[Tuple[int, str]]] = None
    if version >= 2:
        n = c.u32()
        c.fits(n, 8, "regex pattern table")
        regexes, seen = [], set()
        for _ in range(n):
            pid = c.u32()
            if pid >= npat:
                raise IndexFormatError("compiled index regex pattern ID is out of bounds")
            if pid in seen:
                raise IndexFormatError("compiled index contains duplicate regex pattern IDs")
            seen.add(pid)
            try:
                regexes.append((pid, c.lp_bytes().decode("utf-8")))
            except UnicodeDecodeError:
                raise IndexFormatError("compiled index regex pattern is not valid UTF-8")
    ids = c.vec()
    if len(ids) != nlit:
        raise IndexFormatError("compiled index literal ID table is inconsistent")
    nmeta = c.u32()
    c.fits(nmeta, 16, "prefix meta table")
    meta = list(struct.unpack(f"<{4 * nmeta}I", c.take(16 * nmeta)))
    nb = c.u32()
    c.fits(nb, 8, "bucket range table")
    ranges = list(struct.unpack(f"<{2 * nb}I", c.take(8 * nb)))
    ne = c.u32()
    c.fits(ne, 8, "prefilter entry table")
    entries = list(struct.unpack(f"<{2 * ne}I", c.take(8 * ne)))
    for b in range(nb):
        if ranges[2 * b] + ranges[2 * b + 1] > ne:
            raise IndexFormatError(f"compiled index prefilter bucket range #{b} points outside the entry table")
    for st, ln in offsets:
        if st + ln > 0xFFFFFFFF:
            raise IndexFormatError("compiled index literal offset overflows")
        if st + ln > len(packed):
            raise IndexFormatError("compiled index literal offset points outside the packed byte buffer")
    tail_start = c.off
    if version < min(OWN_VERSIONS):
        for _ in range(nrx):
            for _ in range(4):
                c.vec()
            c.take(12)
            c.take(1024)
            c.lp_bytes()
            c.vec()
        if c.remaining():
            raise IndexFormatError("compiled index contains trailing bytes")
    return {
        "version": version, "case_insensitive": ci, "pattern_count": npat, "literal_count": nlit,
        "regex_count": nrx, "hash_window_len": window, "packed": packed, "offsets": offsets,
        "names": names, "regexes": regexes, "literal_ids": ids, "prefix_meta": meta,
        "bucket_ranges": ranges, "entries": entries,
        # Versions 4 and 5: the library's own sections follow here (not restated).
        "tail": data[tail_start:end],
    }


def to_pattern_list(p: dict) -> List[Tuple[int, bytes, Optional[str]]]:
    """The patterns to_pattern_set rebuilds, as (kind, bytes, name)."""
    npat = p["pattern_count"]
    if p["regexes"] is None and p["regex_count"] > 0:
        raise IndexFormatError("compiled index lacks serialized regex patterns")
    rx = {pid: src for pid, src in (p["regexes"] or [])}
    lit = {}
    for li, pid in enumerate(p["literal_ids"]):
        if pid >= npat:
            raise IndexFormatError("compiled index literal pattern ID is out of bounds")
        if pid in lit:
            raise IndexFormatError("compiled index contains duplicate literal IDs")
        lit[pid] = p["offsets"][li]
    out = []
    for pid in range(npat):
        name = p["names"][pid]
        if pid in rx:
            out.append((1, rx[pid].encode(), name))
            continue
        if pid not in lit:
            raise IndexFormatError("compiled index has missing pattern metadata")
        st, ln = lit[pid]
        b = p["packed"][st:st + ln]
        try:
            b.decode("utf-8")
            keep = name
        except UnicodeDecodeError:
            keep = None
        out.append((0, b, keep))
    return out
