"""Compiled pattern index (WPSIDX01) on the host: the oracle's restatement of
the reference format pinned against the reference's own tests and textbook
facts, then the library's writer and reader checked against it.

Reference: core/src/compiled_index/{mod.rs,builder.rs,tests.rs,adversarial_tests/}.
No GPU: ws_index_build, ws_index_inspect and ws_index_save_file are host-only.
"""
import ctypes
import os
import random
import struct
import zlib

import numpy as np
import pytest

import warpstate_amd as w
from warpstate_amd import _native as N
from oracle import index_format as F
from oracle.oracle import fnv1a_window, prefilter_table

# ---------------------------------------------------------------------------
# helpers


def c_patterns(pats):
    bufs = [ctypes.create_string_buffer(b, len(b)) for _, b, _ in pats]
    arr = (N.WsPattern * len(pats))()
    for i, ((k, b, name), buf) in enumerate(zip(pats, bufs)):
        arr[i].kind = k
        arr[i].bytes = ctypes.cast(buf, ctypes.c_void_p)
        arr[i].len = len(b)
        arr[i].name = name.encode() if name is not None else None
    return arr, bufs


def lib_build(pats, ci=False, version=5):
    """ws_index_build: size query, then the write."""
    lib = N.load()
    arr, _keep = c_patterns(pats)
    n = ctypes.c_uint64(0)
    flags = N.INDEX_REFERENCE_V3 if version == 3 else 0
    N.check(lib.ws_index_build(arr, len(pats), 1 if ci else 0, flags, None, 0, ctypes.byref(n)))
    out = ctypes.create_string_buffer(n.value)
    N.check(lib.ws_index_build(arr, len(pats), 1 if ci else 0, flags, out, n.value, ctypes.byref(n)))
    return out.raw[: n.value]


def inspect(data):
    hdr, info = N.WsIndexInfo(), N.WsPatternsInfo()
    st = N.load().ws_index_inspect(bytes(data) if data else None, len(data), ctypes.byref(hdr), ctypes.byref(info))
    msg = N.load().ws_last_error().decode(errors="replace") if st else ""
    return st, msg, {k: getattr(hdr, k) for k, _ in N.WsIndexInfo._fields_}, \
        {k: getattr(info, k) for k, _ in N.WsPatternsInfo._fields_}


def analyze(pats, ci=False):
    arr, _keep = c_patterns(pats)
    info = N.WsPatternsInfo()
    N.check(N.load().ws_patterns_analyze(arr, len(pats), 1 if ci else 0, ctypes.byref(info)))
    return {k: getattr(info, k) for k, _ in N.WsPatternsInfo._fields_}


def recrc(b):
    """Recompute the CRC trailer so that a structural check is what fails."""
    b = bytearray(b)
    b[-4:] = struct.pack("<I", F.crc32_ieee(bytes(b[:-4])))
    return bytes(b)


def lits(*xs):
    return [(0, x if isinstance(x, bytes) else x.encode(), None) for x in xs]


MIXED = [(0, b"alpha", None), (1, b"b[0-9]+", "num"), (0, b"beta", "cred"), (1, b"[A-Z]{3}-[0-9]+", None),
         (0, b"p", None), (0, b"p6", None)]

# ---------------------------------------------------------------------------
# the oracle, pinned


def test_crc32_check_value_and_zlib():
    # CRC-32/ISO-HDLC check value (the IEEE 802.3 CRC of "123456789").
    assert F.crc32_ieee(b"123456789") == 0xCBF43926
    assert F.crc32_ieee(b"") == 0
    rng = random.Random(1)
    for n in (1, 7, 8, 9, 63, 1000):
        b = bytes(rng.randrange(256) for _ in range(n))
        assert F.crc32_ieee(b) == zlib.crc32(b)


def test_oracle_header_layout_matches_reference_tests():
    # adversarial_tests: pattern count at 16, literal count at 20 (load_literal_count_mismatch_detected),
    # header 32 bytes, packed-bytes length at 32 (oversized_packed_bytes_length_detected).
    b = F.build(lits("first", "second", "third"))
    assert b[:8] == b"WPSIDX01"
    assert struct.unpack_from("<7I", b, 8) == (3, 0, 3, 3, 0, 8, len(b"firstsecondthird"))
    assert b[36:36 + 16] == b"firstsecondthird"
    assert struct.unpack_from("<I", b, len(b) - 4)[0] == zlib.crc32(b[:-4])


def test_oracle_prefilter_section_is_the_reference_table():
    ps = lits("alpha", "beta", "a", "abcdefghijk", "zz")
    p = F.parse(F.build(ps))
    lens = [len(x) for _, x, _ in ps]
    hashes = [fnv1a_window(x, 8) for _, x, _ in ps]
    meta, ranges, entries = prefilter_table(lens, hashes, 8)
    assert p["prefix_meta"] == meta.tolist()
    assert p["bucket_ranges"] == ranges.tolist()
    assert p["entries"] == entries.tolist()
    # each literal is found under its own prefix length (literal_prefilter.rs:91-119)
    for i, (_, x, _) in enumerate(ps):
        pl = max(1, min(len(x), 8))
        m = p["prefix_meta"][4 * (pl - 1): 4 * pl]
        b = m[0] + (hashes[i] & m[1])
        st, n = p["bucket_ranges"][2 * b: 2 * b + 2]
        assert i in [p["entries"][2 * e + 1] for e in range(st, st + n) if p["entries"][2 * e] == hashes[i]]


@pytest.mark.parametrize("version", [1, 2, 3])
def test_oracle_round_trip(version):
    pats = MIXED if version >= 2 else lits("a", "bb")
    p = F.parse(F.build(pats, version=version))
    assert p["version"] == version
    assert F.to_pattern_list(p) == pats


def _corruptions(valid):
    """The reference's adversarial cases (adversarial_tests/*.rs) as (name, bytes, keywords)."""
    out = []
    for cut in [8, 9, 12, 16, 20, 24, 28, 31, len(valid) // 2]:
        out.append((f"truncated_{cut}", valid[:cut], ("truncated", "CRC", "invalid")))
    for name, magic in [("garbage", b"NOTINDEX"), ("v00", b"WPSIDX00"), ("v02", b"WPSIDX02"),
                        ("case", b"wpsidx01"), ("ff", b"\xff" * 8), ("zeros", b"\0" * 8)]:
        out.append((name, magic + b"\xff" * (len(valid) - 8), ("header", "CRC", "truncated", "version")))
    out.append(("empty", b"", ("truncated",)))
    v = bytearray(valid)
    v[8:12] = struct.pack("<I", 0)
    out.append(("version_zero", bytes(v), ("version", "invalid")))
    v = bytearray(valid)
    v[8:12] = struct.pack("<I", 6)  # one past the newest version either reader knows
    out.append(("future_version", bytes(v), ("version", "unsupported")))
    v = bytearray(valid)
    v[8:12] = struct.pack("<I", 0xFFFFFFFF)
    out.append(("max_version", bytes(v), ("version", "unsupported")))
    out.append(("trailing", valid + b"TRAILING_GARBAGE_BYTES", ("trailing", "CRC")))
    v = bytearray(valid)
    v[20:24] = struct.pack("<I", 5)
    out.append(("literal_count", recrc(v), ("literal", "inconsistent")))
    v = bytearray(valid)
    v[20:24] = struct.pack("<I", 0)
    out.append(("literal_count_zero", recrc(v), ("literal", "inconsistent")))
    v = bytearray(valid)
    v[16:20] = struct.pack("<I", 100)
    out.append(("pattern_count", recrc(v), ("name", "inconsistent")))
    v = bytearray(valid)
    v[32:36] = struct.pack("<I", 0xFFFFFFFF)
    out.append(("packed_len", recrc(v), ("truncated",)))
    plen = struct.unpack_from("<I", valid, 32)[0]
    off = 36 + plen + 4  # first (start, len) pair
    v = bytearray(valid)
    v[off:off + 8] = struct.pack("<II", 0xFFFFFFFF, 1)
    out.append(("offset_overflow", recrc(v), ("overflows",)))
    v = bytearray(valid)
    v[off:off + 4] = struct.pack("<I", 0xFFFF0000)
    out.append(("offset_past_packed", recrc(v), ("outside",)))
    v = bytearray(valid)
    v[40] ^= 1
    out.append(("bit_flip", bytes(v), ("CRC",)))
    return out


CORRUPTIONS = _corruptions(F.build(lits("first", "second", "third")))


@pytest.mark.parametrize("name,data,keys", CORRUPTIONS, ids=[c[0] for c in CORRUPTIONS])
def test_oracle_rejects_reference_corruptions(name, data, keys):
    with pytest.raises(F.IndexFormatError) as e:
        F.to_pattern_list(F.parse(data))
    assert any(k in str(e.value) for k in keys), str(e.value)


def test_oracle_to_pattern_set_rules():
    # a named literal that is not UTF-8 loses its name (mod.rs:352-363)
    p = F.parse(F.build([(0, b"\xff\xfe", "bin"), (0, b"txt", "t")]))
    assert F.to_pattern_list(p) == [(0, b"\xff\xfe", None), (0, b"txt", "t")]
    # version 1 with regexes stores no sources (mod.rs:287-293)
    with pytest.raises(F.IndexFormatError, match="lacks serialized regex"):
        F.to_pattern_list(F.parse(F.build(MIXED, version=1)))

# ---------------------------------------------------------------------------
# the library's writer


@pytest.mark.parametrize("case", ["basic", "names", "binary", "ci", "single", "many", "long"])
def test_v3_writer_matches_reference_bytes(case):
    rng = random.Random(hash(case) & 0xFFFF)
    ci = case == "ci"
    if case == "basic":
        pats = lits("alpha", "beta", "p", "p6")
    elif case == "names":
        pats = [(0, b"password", "cred"), (0, b"token", None), (0, "café".encode(), "üml")]
    elif case == "binary":
        pats = lits(b"\0abc", b"abc\0", b"a\0b", bytes(range(256)))
    elif case == "ci":
        pats = lits("Needle", "HAY", "stack")
    elif case == "single":
        pats = lits("x")
    elif case == "many":
        pats = lits(*[f"pattern_{i:03}" for i in range(100)])
    else:
        pats = lits(*["".join(rng.choice("ab") for _ in range(rng.randrange(1, 40))) for _ in range(300)])
    assert lib_build(pats, ci, version=3) == F.build(pats, ci, version=3)


def test_regex_that_is_a_literal_is_stored_as_one():
    # compiler.rs:172-211: a regex without metacharacters becomes a literal
    got = lib_build([(1, b"hello", None), (0, b"world", None), (1, rb"a\.b", "dot")], version=3)
    assert got == F.build([(0, b"hello", None), (0, b"world", None), (0, b"a.b", "dot")], version=3)


def test_v3_refuses_regex_sets():
    with pytest.raises(w.WarpstateError) as e:
        lib_build(MIXED, version=3)
    assert e.value.variant == "InvalidArgument"


def test_v5_shares_the_reference_sections():
    for pats, ci in ((MIXED, False), (lits("Needle", "hay"), True), (lits(*[f"k{i}" for i in range(50)]), False)):
        mine = lib_build(pats, ci)
        p = F.parse(mine)
        assert p["version"] == 5
        assert F.to_pattern_list(p) == pats
        ref = F.build(pats, ci, version=3)
        rp = F.parse(ref)
        # everything from the case flag to the end of the prefilter table is the reference's
        assert mine[12:len(mine) - 4 - len(p["tail"])] == ref[12:len(ref) - 4 - len(rp["tail"])]


def test_size_query_and_small_buffer():
    lib = N.load()
    arr, _k = c_patterns(MIXED)
    n = ctypes.c_uint64(0)
    N.check(lib.ws_index_build(arr, len(MIXED), 0, 0, None, 0, ctypes.byref(n)))
    need = n.value
    small = ctypes.create_string_buffer(16)
    st = lib.ws_index_build(arr, len(MIXED), 0, 0, small, 16, ctypes.byref(n))
    assert st == 13 and n.value == need


def test_empty_and_invalid_sets_are_refused():
    lib = N.load()
    n = ctypes.c_uint64(0)
    assert lib.ws_index_build(None, 0, 0, 0, None, 0, ctypes.byref(n)) == 1   # EmptyPatternSet
    with pytest.raises(w.WarpstateError) as e:
        lib_build([(1, b"(a+)+", None)])
    assert e.value.variant == "PathologicalRegex"

# ---------------------------------------------------------------------------
# the library's reader


@pytest.mark.parametrize("version", [1, 2, 3])
def test_reads_reference_files(version):
    pats = MIXED if version >= 2 else lits("alpha", "beta", "gamma")
    st, msg, hdr, info = inspect(F.build(pats, version=version))
    assert st == 0, msg
    assert hdr["version"] == version and hdr["prebuilt"] == 0
    assert hdr["pattern_count"] == len(pats)
    assert info == analyze(pats)


def test_reads_own_files_with_the_same_set_facts():
    for pats, ci in ((MIXED, False), (lits("Needle", "hay"), True)):
        st, msg, hdr, info = inspect(lib_build(pats, ci))
        assert st == 0, msg
        assert hdr["version"] == 5 and hdr["prebuilt"] == 1 and hdr["case_insensitive"] == int(ci)
        assert info == analyze(pats, ci)


def test_version1_regex_file_is_refused():
    st, msg, _, _ = inspect(F.build(MIXED, version=1))
    assert st == 6 and "lacks serialized regex" in msg


@pytest.mark.parametrize("name,data,keys", CORRUPTIONS, ids=[c[0] for c in CORRUPTIONS])
def test_reader_rejects_reference_corruptions(name, data, keys):
    st, msg, _, _ = inspect(data)
    assert st == 6, (st, msg)
    assert any(k in msg for k in keys), msg


def test_v5_corruptions():
    valid = lib_build(MIXED)
    assert inspect(valid)[0] == 0
    v = bytearray(valid)
    v[8:12] = struct.pack("<I", 3)  # a version-5 body is not a version-3 file
    st, msg, _, _ = inspect(recrc(v))
    assert st == 6
    p = F.parse(valid)
    tail0 = len(valid) - 4 - len(p["tail"])
    # first regex section: nstates, nclass, start[4], cmap[256], table
    nst, ncls = struct.unpack_from("<II", valid, tail0)
    for field, value, key in [(0, 1, "state count"), (1, 0, "class count"), (1, 300, "class count"),
                              (2, 0, "start state"), (5, nst, "start state")]:
        v = bytearray(valid)
        v[tail0 + 4 * field: tail0 + 4 * field + 4] = struct.pack("<I", value)
        st, msg, _, _ = inspect(recrc(v))
        assert st == 6 and key in msg, msg
    # a transition to a state that does not exist
    t0 = tail0 + 24 + 256 + 4
    v = bytearray(valid)
    v[t0:t0 + 4] = struct.pack("<I", nst)
    st, msg, _, _ = inspect(recrc(v))
    assert st == 6 and "state range" in msg, msg
    # an end-of-text entry that continues to a state
    e0 = t0 + 4 * ncls
    v = bytearray(valid)
    v[e0:e0 + 4] = struct.pack("<I", 1)
    st, msg, _, _ = inspect(recrc(v))
    assert st == 6 and "end-of-text" in msg, msg
    # a byte class outside the class range
    v = bytearray(valid)
    v[tail0 + 24 + 65] = ncls
    st, msg, _, _ = inspect(recrc(v))
    assert st == 6 and "class range" in msg, msg
    # one regex section too few
    v = bytearray(valid)
    v[24:28] = struct.pack("<I", 1)
    st, msg, _, _ = inspect(recrc(v))
    assert st == 6


def to_version4(v5: bytes) -> bytes:
    """The same set in the earlier version-4 layout of the library's regex
    sections (one start flag, no end-of-text column)."""
    p = F.parse(v5)
    tail0 = len(v5) - 4 - len(p["tail"])
    t, off, out = p["tail"], 0, bytearray()
    for _ in range(p["regex_count"]):
        nst, ncls = struct.unpack_from("<II", t, off)
        cmap = t[off + 24: off + 280]
        n = struct.unpack_from("<I", t, off + 280)[0]
        table = np.frombuffer(t, dtype="<u4", count=n, offset=off + 284).reshape(nst, ncls + 1)[:, :ncls]
        out += struct.pack("<III", nst, ncls, 0) + cmap + struct.pack("<I", nst * ncls) + table.astype("<u4").tobytes()
        off += 284 + 4 * n
    out += t[off:]  # the factor automaton section
    body = bytearray(v5[:tail0]) + out
    body[8:12] = struct.pack("<I", 4)
    return bytes(body) + struct.pack("<I", zlib.crc32(bytes(body)) & 0xFFFFFFFF)


def test_version4_files_are_rebuilt_from_their_sources():
    v4 = to_version4(lib_build(MIXED))
    st, msg, hdr, info = inspect(v4)
    assert st == 0, msg
    assert hdr["version"] == 4 and hdr["prebuilt"] == 0
    assert info == analyze(MIXED)


def test_every_truncation_is_an_error():
    for data in (F.build(lits("ab", "cd")), lib_build(MIXED)):
        for cut in range(len(data)):
            st, _msg, _, _ = inspect(data[:cut])
            assert st == 6, cut


def test_random_damage_never_passes_silently_wrong_structure():
    # Flip bytes and re-seal the CRC: every outcome is a clean accept or a
    # PatternCompilationFailed-class error, never a crash.
    valid = lib_build(MIXED)
    rng = random.Random(7)
    outcomes = set()
    for _ in range(1500):
        v = bytearray(valid)
        for _ in range(rng.randrange(1, 4)):
            v[rng.randrange(len(v) - 4)] = rng.randrange(256)
        st, _msg, _, _ = inspect(recrc(v))
        assert st in (0, 2, 6, 9, 12), st
        outcomes.add(st)
    assert 6 in outcomes


def test_save_and_load_file_round_trip(tmp_path):
    ps = w.PatternSet.builder().literal("disk").regex("lo+ad").build()
    ix = w.CompiledPatternIndex.load(w.CompiledPatternIndex.build(ps))
    path = tmp_path / "set.idx"
    ix.save_to_file(path)
    again = w.CompiledPatternIndex.load_from_file(path)
    assert again.to_bytes() == ix.to_bytes() and again.prebuilt and again.literal_count() == 1


def test_load_from_file_size_limit(tmp_path):
    path = tmp_path / "huge.idx"
    with open(path, "wb") as f:
        f.truncate(512 * 1024 * 1024 + 1)  # sparse
    with pytest.raises(w.WarpstateError) as e:
        w.CompiledPatternIndex.load_from_file(path)
    assert e.value.variant == "InputTooLarge"


def test_load_is_faster_than_compiling():
    # The point of version 5: rebuilding from the stored tables skips the
    # subset construction and the automaton build.
    import time

    from warpstate_amd import workloads as wl

    pats = [(1, s.source.encode(), None) for s in wl.regex_set(60, seed=5)]
    data = lib_build(pats)
    t0 = time.perf_counter()
    for _ in range(3):
        assert inspect(data)[0] == 0
    t1 = time.perf_counter()
    for _ in range(3):
        analyze(pats)
    t2 = time.perf_counter()
    assert (t1 - t0) < 0.5 * (t2 - t1), (t1 - t0, t2 - t1)
