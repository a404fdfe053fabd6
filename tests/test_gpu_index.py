"""Compiled pattern index on the device: a set loaded from a file (version 5
with its stored tables, version 3 written here, versions 1-3 written by the
oracle's restatement of the reference builder) scans exactly like the set
compiled from its patterns, and like the oracle.

Reference: core/src/compiled_index/tests.rs:5-128 (round trips), mod.rs:274-276
(names), adversarial_tests/proptest_roundtrip_preserves_scan_results.rs.
"""
import ctypes
import random

import numpy as np
import pytest

import warpstate_amd as w
from oracle import index_format as F
from oracle.oracle import Oracle
from warpstate_amd import _native as N
from warpstate_amd import workloads as W

pytestmark = pytest.mark.gpu


def build(pats, ci=False):
    b = w.PatternSet.builder().case_insensitive(ci)
    for kind, p, name in pats:
        if kind == 0:
            b = b.named_literal(name, p) if name else b.literal(p)
        else:
            b = b.named_regex(name, p) if name else b.regex(p)
    return b.build()


def both_modes(ps, data):
    m = w.GpuMatcher(ps)
    return m.scan(data, mode="overlapping"), m.scan(data, mode="standard")


def mixed_case(seed):
    pats, lits, specs = W.mixed_set(60, 20, seed=seed)
    data = bytearray(W.printable_input(1 << 18, seed=seed).tobytes())
    rng = random.Random(seed)
    nrng = np.random.default_rng(seed)
    pieces = list(lits) + [s.sample(nrng) for s in specs for _ in range(3)]
    for p in pieces:
        o = rng.randrange(len(data) - len(p))
        data[o:o + len(p)] = p
    return [(k, p, None) for k, p in pats], bytes(data)


def test_reference_round_trip_cases():
    # tests.rs:5-16, 19-31, 88-98
    ps = build([(0, b"alpha", None), (1, b"b[0-9]+", None), (0, b"beta", None)])
    ix = w.CompiledPatternIndex.load(w.CompiledPatternIndex.build(ps))
    assert ix.prebuilt and ix.scan(b"alpha b42 beta") == ps.scan(b"alpha b42 beta")
    assert [tuple(m) for m in ix.scan(b"alpha b42 beta")] == [(0, 0, 5), (1, 6, 9), (2, 10, 14)]
    ps = build([(0, b"Needle", None)], ci=True)
    ix = w.CompiledPatternIndex.load(w.CompiledPatternIndex.build(ps))
    assert ix.scan(b"xxneedlexx") == ps.scan(b"xxneedlexx") == [(0, 2, 8)]
    ps = build([(0, f"pattern_{i:03}".encode(), None) for i in range(100)])
    for version in (3, 5):
        ix = w.CompiledPatternIndex.load(w.CompiledPatternIndex.build(ps, version=version))
        got = ix.scan(b"xxpattern_042xxpattern_099xx")
        assert got == ps.scan(b"xxpattern_042xxpattern_099xx") == [(42, 2, 13), (99, 15, 26)]


@pytest.mark.parametrize("seed", [11, 12])
def test_loaded_set_scans_like_compiled_set(seed):
    pats, data = mixed_case(seed)
    ps = build(pats)
    want = both_modes(ps, data)
    orc = Oracle([(k, p) for k, p, _ in pats])
    assert [tuple(m) for m in want[0]] == orc.scan_list(data, 0)
    assert [tuple(m) for m in want[1]] == orc.scan_list(data, 1)
    ix = w.CompiledPatternIndex.load(w.CompiledPatternIndex.build(ps))
    assert ix.prebuilt
    assert both_modes(ix.to_pattern_set(), data) == want
    assert ix.to_pattern_set().info() == ps.info()


def test_kmer_set_round_trip_at_size():
    lits = W.kmer_set(2000, seed=3)
    data = W.kmer_input(1 << 22, lits, seed=3)
    ps = build([(0, p, None) for p in lits])
    ix = w.CompiledPatternIndex.load(w.CompiledPatternIndex.build(ps))
    m0, m1 = w.GpuMatcher(ps), w.GpuMatcher(ix.to_pattern_set())
    a, b = m0.scan_array(data.tobytes()), m1.scan_array(data.tobytes())
    assert len(a) > 100 and np.array_equal(a, b)
    # property: every reported match is the literal it names
    raw = data.tobytes()
    for pid, s, e in a[:: max(1, len(a) // 500)]:
        assert raw[s:e] == lits[pid]


@pytest.mark.parametrize("version", [1, 2, 3])
def test_reference_written_files_load_and_scan(version):
    pats, data = mixed_case(21)
    if version == 1:
        pats = [p for p in pats if p[0] == 0]
    ref = F.build(pats, version=version)
    ix = w.CompiledPatternIndex.load(ref)
    assert ix.version == version and not ix.prebuilt
    assert both_modes(ix.to_pattern_set(), data) == both_modes(build(pats), data)


def test_case_insensitive_set_and_names():
    pats = [(0, b"Password", "cred"), (1, b"[A-Z]{4}-[0-9]{4}", "key"), (0, b"TOKEN", None)]
    ps = build(pats, ci=True)
    data = b"password ABCD-1234 token PASSWORD abcd-1234"
    ix = w.CompiledPatternIndex.load(w.CompiledPatternIndex.build(ps))
    assert ix.scan(data) == ps.scan(data)
    assert ix.names() == ["cred", "key", None]
    # named_patterns_survive_round_trip (tests.rs:100-113)
    ps2 = build([(0, b"password", "cred"), (1, b"[A-Z]{4}-[0-9]{4}", "key")])
    ix2 = w.CompiledPatternIndex.load(w.CompiledPatternIndex.build(ps2))
    assert len(ix2.scan(b"password ABCD-1234")) == 2
    # a named literal that is not UTF-8 is rebuilt without its name (mod.rs:352-363)
    ix3 = w.CompiledPatternIndex.load(F.build([(0, b"\xff\xfe", "bin"), (0, b"txt", "t")]))
    assert ix3.names() == [None, "t"]
    assert ix3.scan(b"a\xff\xfetxt") == [(0, 1, 3), (1, 3, 6)]


def test_serialize_of_a_loaded_set_is_the_same_file():
    pats, _ = mixed_case(31)
    b4 = w.CompiledPatternIndex.build(build(pats))
    ix = w.CompiledPatternIndex.load(b4)
    assert w.CompiledPatternIndex.build(ix.to_pattern_set()) == b4
    lits = [(0, f"lit{i}".encode(), None) for i in range(40)]
    b3 = F.build(lits)
    ix3 = w.CompiledPatternIndex.load(b3)
    assert w.CompiledPatternIndex.build(ix3.to_pattern_set(), version=3) == b3


def test_c_abi_file_entry_points(tmp_path):
    lib = N.load()
    ctx = w.default_context(0)
    ps = build([(0, b"disk", None), (1, b"lo+ad", None)])
    data = w.CompiledPatternIndex.build(ps)
    path = str(tmp_path / "set.idx").encode()
    N.check(lib.ws_index_save_file(data, len(data), path))
    h = ctypes.c_void_p()
    N.check(lib.ws_index_load_file(ctx.handle, path, ctypes.byref(h)))
    try:
        out = np.zeros(16, dtype=N.MATCH_DTYPE)
        cnt = ctypes.c_uint64(0)
        src = np.frombuffer(b"disk loooad", dtype=np.uint8)
        N.check(lib.ws_scan(ctx.handle, h, src.ctypes.data, len(src), N.MEM_HOST, N.MODE_STANDARD, None,
                            out.ctypes.data, 16, N.MEM_HOST, ctypes.byref(cnt)))
        assert [tuple(m) for m in out[: cnt.value]] == [(0, 0, 4), (1, 5, 11)]
    finally:
        lib.ws_patterns_free(h)
    st = lib.ws_index_load_file(ctx.handle, str(tmp_path / "missing.idx").encode(), ctypes.byref(h))
    assert st == 6 and not h.value
    bad = bytearray(data)
    bad[40] ^= 1
    st = lib.ws_index_load(ctx.handle, bytes(bad), len(bad), ctypes.byref(h))
    assert st == 6 and not h.value and b"CRC" in lib.ws_last_error()


def test_random_literal_sets_round_trip():
    # proptest_roundtrip_preserves_scan_results: [a-zA-Z0-9]{1,20} x 1..50, input [a-zA-Z0-9 ]{0,1000}
    rng = random.Random(5)
    alpha = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789"
    for case in range(25):
        pats = [(0, "".join(rng.choice(alpha) for _ in range(rng.randint(1, 20))).encode(), None)
                for _ in range(rng.randint(1, 50))]
        text = "".join(rng.choice(alpha + " ") for _ in range(rng.randint(0, 1000))).encode()
        ps = build(pats)
        ix = w.CompiledPatternIndex.load(w.CompiledPatternIndex.build(ps, version=3 if case % 2 else 5))
        got = ix.scan(text)
        assert got == ps.scan(text)
        assert [tuple(m) for m in got] == Oracle([(0, p) for _, p, _ in pats]).scan_list(text, 1)
