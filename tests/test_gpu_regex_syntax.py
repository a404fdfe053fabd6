"""GPU parity for the regex syntax beyond literal-prefix patterns: `^` `$`
(multi-line), `\\A` `\\z`, `\\b` `\\B`, regexes that can match the empty
string (the regex crate's find_iter rule for empty matches) and regexes
whose first byte is unselective (every position a candidate, k_rx_scan).

Against the committed vectors and KATs (Python `re` driven by the regex
crate's iteration rule; the reference's own assertions), and against the
oracle on seeded inputs: whole scans on every kernel variant and lane width,
both pipelines, streams (the byte before each chunk decides `^` and `\\b`),
batches (each item its own text, empty items included), empty inputs and
the allocation-free entry point.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from warpstate_amd import (GpuMatcher, PatternSet, ScanItem, StreamPipeline, Tuning, VARIANT_DFA_GLOBAL,
                           VARIANT_DFA_LDS, VARIANT_HASH, scan_batch, scan_batch_device)
from warpstate_amd import _native as N
from warpstate_amd import workloads as W
from warpstate_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))

LOOK_SET = [r"^abc", r"end$", r"\bfoo\b", r"\Bo\B", r"^[a-z]+$", r"[^a-z]", r"x?", r"(?s).{2}x",
            r"[^\n]{1,5}end", r"\Aab", r"ab\z", r"^$", r"[0-9]*\b", r"start|^fin"]


def build(pats):
    b = PatternSet.builder()
    for kind, p in pats:
        (b.literal if kind == 0 else b.regex)(p)
    return b.build()


def as_list(arr):
    return [(int(p), int(s), int(e)) for p, s, e in zip(arr["pattern_id"], arr["start"], arr["end"])]


def text(size, seed, line=40):
    r = np.random.default_rng(seed)
    alpha = np.frombuffer(b"abcfooendstartxyq0129_ -AB", dtype=np.uint8)
    d = alpha[r.integers(0, len(alpha), size=size)]
    d[r.integers(0, size, size=max(1, size // line))] = ord("\n")
    return d.tobytes()


def load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


def test_regex_vectors_on_device():
    vecs = load("regex_vectors.json")
    for k, v in enumerate(vecs):
        data = bytes.fromhex(v["input"])
        m = GpuMatcher(build([(1, v["pattern"])]))
        exp = [(0, s, e) for s, e in v["expected"]]
        for t in (None, Tuning(checked=True)):
            got = as_list(m.scan_array(data, tuning=t))
            assert got == exp, (k, v["pattern"], t, got[:4], exp[:4])


def test_reference_kats_on_device():
    for kat in load("kats.json"):
        pats = [(kind, bytes.fromhex(h)) for kind, h in kat["patterns"]]
        b = PatternSet.builder().case_insensitive(kat["case_insensitive"])
        for kind, p in pats:
            (b.literal_bytes if kind == 0 else b.regex)(p if kind == 0 else p.decode("latin-1"))
        m = GpuMatcher(b.build())
        mode = "overlapping" if kat["mode"] == "overlapping" else "standard"
        got = as_list(m.scan_array(bytes.fromhex(kat["input"]), mode=mode))
        assert got == [tuple(e) for e in kat["expected"]], (kat["name"], kat["source"], got)


def test_look_set_every_variant_and_lane_width():
    pats = [(1, p) for p in LOOK_SET] + [(0, b"start"), (0, b"xy")]
    pats += [(1, s.source) for s in W.regex_set(20, seed=31)]
    data = text(300_000, 32)
    m = GpuMatcher(build(pats))
    exp = Oracle(pats).scan_list(data)
    assert len(exp) > 10_000
    v = m.info()["variants_supported"]
    tunings = [Tuning(variant=VARIANT_HASH, bytes_per_lane=b) for b in (16, 32, 64)] + [Tuning(checked=True)]
    tunings += [Tuning(variant=x) for x in (VARIANT_DFA_GLOBAL, VARIANT_DFA_LDS) if v & (1 << x)]
    for t in tunings:
        print(f"look set tuning={t}", flush=True)
        got = as_list(m.scan_array(data, tuning=t))
        assert got == exp, (t, len(got), len(exp))
    # the standard mode adds nothing for regexes: same regex matches
    assert [x for x in as_list(m.scan_array(data, mode="standard")) if x[0] < len(LOOK_SET)] == \
           [x for x in exp if x[0] < len(LOOK_SET)]


@pytest.mark.parametrize("seed", range(3))
def test_look_grammar_sets_against_oracle(seed):
    from test_oracle import _look_regex

    r = np.random.default_rng(1300 + seed)
    pats = [(1, _look_regex(r)) for _ in range(25)]
    data = text(120_000, 1400 + seed, line=30)
    exp = Oracle(pats).scan_list(data)
    got = as_list(GpuMatcher(build(pats)).scan_array(data))
    assert got == exp, (len(got), len(exp), [p for _, p in pats])


@pytest.mark.parametrize("chunk", [4096, 65_536, 100_001])
def test_streams_see_the_byte_before_each_chunk(chunk):
    # Chunks start mid-line, right after a newline and inside a word: `^`,
    # `\b` and `\B` at a chunk's first byte read the byte before it.
    pats = [(1, r"^start"), (1, r"\bfoo\b"), (1, r"\Bo"), (1, r"[^a-z]{1,2}"), (1, r"x?"), (1, r"end$")]
    data = bytearray(text(400_000, 41))
    for pos in range(chunk - 1, len(data), chunk):
        data[pos] = ord("\n")  # the next chunk starts a line
    data = bytes(data)
    m = GpuMatcher(build(pats))
    exp = Oracle(pats).scan_list(data)
    whole = as_list(m.scan_array(data))
    assert whole == exp
    sp = StreamPipeline(m, 8)
    for t in (None, Tuning(checked=True)):
        assert as_list(sp.scan_array(data, chunk_size=chunk, tuning=t)) == exp, t
    dev = DeviceBuffer.from_array(np.frombuffer(data, dtype=np.uint8))
    out = DeviceBuffer(12 * (len(exp) + 64))
    for k in range(3):  # eager, captured, replayed
        n = sp.scan_device(dev.ptr, len(data), out.ptr, len(exp) + 64, chunk_size=chunk)
        got = out.download(np.uint8, 12 * n).view(N.MATCH_DTYPE)
        assert as_list(got) == exp, k


def test_batch_items_are_their_own_texts():
    # `^`, `\A` and `\b` hold at each item's start, `$` and `\z` at its end;
    # an empty item has the empty matches of an empty text.
    pats = [(1, r"^ab"), (1, r"\Aab"), (1, r"ab\z"), (1, r"b$"), (1, r"a*"), (1, r"\bab"), (0, b"ab")]
    rng = np.random.default_rng(51)
    items = []
    for i in range(400):
        n = int(rng.integers(0, 40)) if i % 7 else 0
        items.append(ScanItem(i, bytes(rng.choice(np.frombuffer(b"ab \nc", dtype=np.uint8), size=n))))
    m = GpuMatcher(build(pats))
    o = Oracle(pats)
    exp = [(it.id, x) for it in items for x in o.scan_list(it.data)]
    got = [(t.source_id, tuple(t.matched)) for t in scan_batch(m, items)]
    assert got == exp
    # device items
    buf = np.frombuffer(b"".join(it.data for it in items) + b"\0", dtype=np.uint8)
    offs = np.cumsum([0] + [len(it.data) for it in items[:-1]]).astype(np.int64)
    dev = DeviceBuffer.from_array(buf)
    ids = np.arange(len(items), dtype=np.uint64)
    sizes = np.array([len(it.data) for it in items], dtype=np.int64)
    got2 = scan_batch_device(m, ids, (dev.ptr + offs).astype(np.uint64), sizes)
    assert [(int(a), (int(b), int(c), int(d))) for a, b, c, d in
            zip(got2["source_id"], got2["pattern_id"], got2["start"], got2["end"])] == exp


def test_empty_inputs():
    # core/tests/adversarial/break_it.rs:22-27, core/tests/unit/mass_pattern_samples.rs:866-875
    for p in (".*", "a*", "^", "$", r"\A", r"\z", r"\B"):
        m = GpuMatcher(build([(1, p), (0, b"test")]))
        assert m.scan(b"") == [(0, 0, 0)], p
        assert StreamPipeline(m, 8).scan(b"", chunk_size=4096) == [(0, 0, 0)], p
    assert GpuMatcher(build([(1, r"\b"), (0, b"test")])).scan(b"") == []


def test_unselective_regex_overflows_to_the_checked_pipeline():
    # A match of up to 31 bytes from every byte outside [a-z] (about a third
    # of the positions): 4.4 M raw matches, more than a device-sized run of
    # 12 MiB holds, so the call is redone on the checked pipeline; the chain
    # keeps about one in 31.
    pats = [(1, r"(?s)[^a-z].{0,30}"), (1, r"[0-9]{2}\b")]
    data = text(12 << 20, 61)
    m = GpuMatcher(build(pats))
    exp = Oracle(pats).scan(data, 0)
    assert 300_000 < len(exp) < 1_000_000
    got = m.scan_array(data)
    assert np.array_equal(got, exp)


def test_allocation_free_scan_with_look_regexes():
    # Regexes with assertions and unselective first bytes whose matches are
    # sparse enough for the fixed buffers.
    pats = [(1, p) for p in (r"^abc", r"end$", r"\bfoo\b", r"(?s).{2}x", r"[^\n]{1,5}end", r"\Aab",
                             r"ab\z")] + [(0, b"foo")]
    data = text(200_000, 71)
    m = GpuMatcher(build(pats))
    exp = m.scan_array(data)
    assert np.array_equal(exp, Oracle(pats).scan(data, 0))
    dev = DeviceBuffer.from_array(np.frombuffer(data, dtype=np.uint8))
    out = DeviceBuffer(12 * (len(exp) + 16))
    count = ctypes.c_uint64(0)
    st = N.load().ws_scan_noalloc(m.ctx.handle, m.handle, ctypes.c_void_p(dev.ptr), len(data), None,
                                  ctypes.c_void_p(out.ptr), len(exp) + 16, ctypes.byref(count))
    assert st == 0, N.load().ws_last_error()
    assert np.array_equal(out.download(np.uint8, 12 * count.value).view(exp.dtype), exp)
