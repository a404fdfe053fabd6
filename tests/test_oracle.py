"""Pins the CPU oracle before it is trusted as the GPU checker.

* Reference known-answer cases (tests/golden/kats.json, each citing the
  reference test it comes from).
* Python `re` (leftmost-first find_iter) on committed vectors and on fresh
  seeded inputs of the regex grammar.
* Brute force for overlapping literal matches (the definition itself).
* The reference's FNV window hash and prefilter-table KATs
  (core/src/pattern/mod.rs:427-477).
* Invariants of the non-overlapping reductions, chosen so that a wrong
  tie-break or a skipped candidate fails.
"""
import json
import os
import re

import numpy as np
import pytest

from oracle.oracle import Oracle, OracleError, fnv1a_window, prefilter_probe, prefilter_table
from warpstate_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


@pytest.mark.parametrize("kat", load("kats.json"), ids=lambda k: k["name"])
def test_reference_kats(kat):
    pats = [(k, bytes.fromhex(h)) for k, h in kat["patterns"]]
    o = Oracle(pats, kat["case_insensitive"])
    got = o.scan_list(bytes.fromhex(kat["input"]), 0 if kat["mode"] == "overlapping" else 1)
    assert got == [tuple(e) for e in kat["expected"]], kat["source"]


@pytest.mark.parametrize("vec", load("regex_vectors.json"), ids=lambda v: v["pattern"][:24])
def test_regex_vectors(vec):
    o = Oracle([(1, vec["pattern"])])
    assert o.find_iter(0, bytes.fromhex(vec["input"])) == [tuple(e) for e in vec["expected"]]


@pytest.mark.parametrize("seed", range(6))
def test_regex_grammar_vs_python_re(seed):
    specs = W.regex_set(25, seed=100 + seed)
    data = W.regex_input(60_000, specs, seed=seed, plant_gap=700).tobytes()
    o = Oracle([(1, s.source) for s in specs])
    for i, s in enumerate(specs):
        exp = [(m.start(), m.end()) for m in re.finditer(s.source.encode(), data)]
        assert o.find_iter(i, data) == exp, s.source


def test_word_boundary_definitions_at_the_text_ends():
    # The regex crate's ASCII \b / \B: a word byte on exactly one side / on
    # both or neither side, the text's ends counting as non-word. In an empty
    # text \B therefore holds at 0 and \b does not (CPython's `re` refuses
    # \B there, so this case is pinned by the definition).
    assert Oracle([(1, r"\B")]).find_iter(0, b"") == [(0, 0)]
    assert Oracle([(1, r"\b")]).find_iter(0, b"") == []
    assert Oracle([(1, r"\b")]).find_iter(0, b"ab") == [(0, 0), (2, 2)]
    assert Oracle([(1, r"\B")]).find_iter(0, b"ab") == [(1, 1)]
    assert Oracle([(1, r"\B")]).find_iter(0, b" ") == [(0, 0), (1, 1)]


def _look_regex(r):
    """A random regex over assertions, classes and (unnested) repeats."""
    atoms = ["a", "b", "f", "o", "[ab]", "[^ab\n]", ".", r"\w", r"\d", "(?:fo|b)", "x"]
    looks = ["^", "$", r"\b", r"\B", r"\A", r"\z"]
    quants = ["", "", "*", "+", "?", "{0,2}", "{1,3}", "*?", "+?"]
    branches = []
    for _ in range(int(r.integers(1, 4))):
        parts = []
        for _ in range(int(r.integers(1, 5))):
            if r.random() < 0.3:
                parts.append(looks[int(r.integers(0, len(looks)))])
            else:
                parts.append(atoms[int(r.integers(0, len(atoms)))] + quants[int(r.integers(0, len(quants)))])
        branches.append("".join(parts))
    return "|".join(branches)


@pytest.mark.parametrize("seed", range(4))
def test_look_and_empty_grammar_vs_python_re(seed):
    # Seeded regexes with assertions and empty-capable repeats against `re`
    # searches driven by the regex crate's iteration rule.
    from golden.make_golden import py_pattern, rust_find_iter

    r = np.random.default_rng(900 + seed)
    alpha = np.frombuffer(b"abfox1 \n_-", dtype=np.uint8)
    for k in range(60):
        pat = _look_regex(r)
        data = alpha[r.integers(0, len(alpha), size=int(r.integers(1, 300)))].tobytes()
        rx = re.compile(py_pattern(pat), re.M)
        assert Oracle([(1, pat)]).find_iter(0, data) == rust_find_iter(rx, data), (pat, data)


def brute_overlapping(lits, data, ci=False):
    f = (lambda b: b.lower()) if ci else (lambda b: b)
    d = f(data)
    out = []
    for pos in range(len(d)):
        for pid, l in enumerate(lits):
            if d[pos:pos + len(l)] == f(l):
                out.append((pid, pos, pos + len(l)))
    return out


@pytest.mark.parametrize("seed", range(5))
def test_overlapping_equals_brute_force(seed):
    r = np.random.default_rng(seed)
    alpha = np.frombuffer(b"abcA", dtype=np.uint8)
    lits = [alpha[r.integers(0, 4, size=int(r.integers(1, 6)))].tobytes() for _ in range(12)]
    lits.append(lits[0])  # duplicate pattern
    data = alpha[r.integers(0, 4, size=3000)].tobytes()
    for ci in (False, True):
        assert Oracle([(0, l) for l in lits], ci).scan_list(data) == brute_overlapping(lits, data, ci)


def test_fnv_window_kats():
    # core/src/pattern/mod.rs:427-452
    assert fnv1a_window(b"hello", 8) == fnv1a_window(b"hello", 8)
    assert fnv1a_window(b"hello", 8) != fnv1a_window(b"world", 8)
    assert fnv1a_window(b"longer_pattern", 4) == fnv1a_window(b"long", 4)
    assert fnv1a_window(b"ab", 10) == fnv1a_window(b"ab", 2)
    # closed form of FNV-1a-32 (offset basis 2166136261, prime 16777619)
    h = 2166136261
    for b in b"abc":
        h = ((h ^ b) * 16777619) & 0xFFFFFFFF
    assert fnv1a_window(b"abc", 8) == h
    assert fnv1a_window(b"", 8) == 2166136261


def test_prefilter_table_probes_every_literal():
    # core/src/pattern/mod.rs:455-477, plus a large random set
    for lits in ([b"a", b"ab", b"alphabet", b"beta", b"betamax"],
                 W.random_literals(3000, 1, 20, W.LIT_ALPHABET, seed=5)):
        lens = [len(l) for l in lits]
        hashes = [fnv1a_window(l, 8) for l in lits]
        meta, ranges, entries = prefilter_table(lens, hashes, 8)
        for i, (l, h) in enumerate(zip(lits, hashes)):
            assert i in prefilter_probe(meta, ranges, entries, 8, max(1, min(len(l), 8)), h)
        # bucket_count = next_pow2(2 * group size) per prefix length (literal_prefilter.rs:55-59)
        for g in range(1, 9):
            cnt = sum(1 for l in lits if min(len(l), 8) == g)
            if cnt:
                bc = 1
                while bc < 2 * cnt:
                    bc <<= 1
                assert meta[4 * (g - 1) + 2] == bc and meta[4 * (g - 1) + 1] == bc - 1


def non_overlapping(ms):
    ends = {}
    for pid, s, e in ms:
        assert s >= ends.get(pid, 0)
        ends[pid] = e


@pytest.mark.parametrize("seed", range(4))
def test_leftmost_longest_invariants(seed):
    r = np.random.default_rng(seed)
    alpha = np.frombuffer(b"ab", dtype=np.uint8)
    lits = list({alpha[r.integers(0, 2, size=int(r.integers(2, 7)))].tobytes() for _ in range(30)})
    lits = lits[:20] + [lits[0]]  # keep a duplicate: ties go to the earlier id
    data = alpha[r.integers(0, 2, size=4000)].tobytes()
    o = Oracle([(0, l) for l in lits])
    assert o.strategy == 1
    std = o.scan_list(data, 1)
    ov = o.scan_list(data, 0)
    # kept matches never overlap each other (single greedy chain over all ids)
    for a, b in zip(std, std[1:]):
        assert a[2] <= b[1]
    # every kept match is the longest at its start, lowest id among equals
    by_start = {}
    for pid, s, e in ov:
        best = by_start.get(s)
        if best is None or e > best[1] or (e == best[1] and pid < best[0]):
            by_start[s] = (pid, e)
    for pid, s, e in std:
        assert by_start[s] == (pid, e)
    # no overlapping-list start is skipped unless covered by a kept match
    covered = set()
    for _, s, e in std:
        covered.update(range(s, e))
    kept_starts = {s for _, s, _ in std}
    for s in by_start:
        assert s in kept_starts or s in covered


def test_hash_scanner_invariants():
    r = np.random.default_rng(3)
    alpha = np.frombuffer(b"ab", dtype=np.uint8)
    lits = W.random_literals(5200, 6, 12, alpha, seed=3)
    data = alpha[r.integers(0, 2, size=5000)].tobytes()
    o = Oracle([(0, l) for l in lits])
    assert o.strategy == 3
    std = o.scan_list(data, 1)
    ov = o.scan_list(data, 0)
    starts = sorted({s for _, s, _ in std})
    shortest = {}
    for pid, s, e in ov:
        shortest[s] = min(shortest.get(s, 1 << 30), e - s)
    # every id with the shortest length at a kept start is reported (duplicates too)
    for s in starts:
        want = sorted((pid, s, s + shortest[s]) for pid, s2, e in ov if s2 == s and e - s2 == shortest[s])
        assert [m for m in std if m[1] == s] == want
    # the cursor jumps to the end of the shortest match
    for a, b in zip(starts, starts[1:]):
        assert b >= a + shortest[a]
        assert all(x not in shortest for x in range(a + shortest[a], b))


def test_multi_memchr_differs_from_leftmost_longest_where_reference_does():
    # Per-needle non-overlapping occurrences (memmem) before the greedy merge
    # (core/src/specialize.rs:313-373): "aa" at 1 is hidden behind "ba" at 0.
    o = Oracle([(0, b"aa"), (0, b"ba")])
    assert o.strategy == 2
    assert o.scan_list(b"baaa", 1) == [(1, 0, 2)]
    o2 = Oracle([(0, b"aa"), (0, b"ba")] + [(0, b"zzzzzzzzzzzzzzzzzzzz")])  # >16 bytes -> LeftmostLongest
    assert o2.strategy == 1
    assert o2.scan_list(b"baaa", 1) == [(1, 0, 2), (0, 2, 4)]


@pytest.mark.parametrize("pat,status", [("(a+)+", 12), ("([0-9]+)*", 12), ("a(", 6), ("a{3", 6), ("^*", 6)])
def test_compile_errors(pat, status):
    with pytest.raises(OracleError) as e:
        Oracle([(1, pat)])
    assert e.value.status == status


def test_literal_regex_becomes_literal():
    # core/src/pattern/compiler.rs:179-212: a pure-literal regex scans as a literal
    o = Oracle([(1, "abc")])
    assert o.scan_list(b"abcabc", 0) == [(0, 0, 3), (0, 3, 6)]
    o = Oracle([(1, "aa")])
    assert o.scan_list(b"aaa", 0) == [(0, 0, 2), (0, 1, 3)]


def test_threaded_overlapping_scan_equals_single_thread():
    # The byte-range split with a (max_len - 1) halo must reproduce the
    # one-thread overlapping scan exactly, including matches that straddle
    # shard edges (short input, many shards).
    import numpy as np
    from oracle.oracle import MATCH_DTYPE, Oracle

    lits = [b"abcab", b"bca", b"ab", b"cabcabc", b"a"]
    data = (b"abcabcabcab" * 200)[:1999]
    o = Oracle([(0, l) for l in lits])
    ref = o.scan(data, 0)
    for threads in (1, 2, 3, 7, 64):
        out = np.zeros(len(ref) + 8, dtype=MATCH_DTYPE)
        n = o.scan_into(data, out, 0, threads)
        assert n == len(ref)
        assert np.array_equal(out[:n], ref), threads


@pytest.mark.parametrize("mode", [0, 1])
def test_threaded_whole_set_scan_equals_single_thread(mode):
    # or_scan_mt (literal shards or one literal job, one job per regex,
    # merged and sorted) must equal or_scan on a mixed set in both modes.
    pats, _, specs = W.mixed_set(60, 25, seed=21)
    rng = np.random.default_rng(22)
    data = W.printable_input(300_000, 23)
    W.plant(data, [p for k, p in pats if k == 0] + [s.sample(rng) for s in specs], 300, 24)
    o = Oracle(pats)
    ref = o.scan(data, mode)
    assert len(ref) > 500
    for threads in (1, 3, 16):
        assert np.array_equal(o.scan_mt(data, mode, threads), ref), threads


# --- StreamScanner (core/src/stream.rs) -------------------------------------

from oracle.oracle import StreamScannerOracle  # noqa: E402


def _keep(pats):
    return max((len(p) for k, p in pats if k == 0), default=0)


@pytest.mark.parametrize("kat", load("stream_kats.json"), ids=lambda k: k["name"])
def test_stream_scanner_kats(kat):
    pats = [(k, bytes.fromhex(h)) for k, h in kat["patterns"]]
    sc = StreamScannerOracle(Oracle(pats), _keep(pats), mode=1)
    for feed, exp in zip(kat["feeds"], kat["expected"]):
        assert sc.feed(bytes.fromhex(feed)) == [tuple(e) for e in exp], kat["source"]


@pytest.mark.parametrize("seed", range(6))
def test_stream_scanner_overlapping_literals_equal_whole_scan(seed):
    # Every occurrence of a literal ends past the tail of exactly one window
    # (the tail is the longest literal), so the feeds add up to the whole
    # overlapping scan, which is brute force here.
    r = np.random.default_rng(seed)
    pats = [(0, bytes(r.integers(97, 100, size=int(r.integers(1, 6)), dtype=np.uint8))) for _ in range(5)]
    data = bytes(r.integers(97, 100, size=3000, dtype=np.uint8))
    cuts = sorted(r.integers(0, len(data), size=12).tolist())
    sc = StreamScannerOracle(Oracle(pats), _keep(pats), mode=0)
    got = []
    prev = 0
    for c in cuts + [len(data)]:
        got += sc.feed(data[prev:c])
        prev = c
    got += sc.finish()
    brute = sorted((i, s, s + len(p)) for i, (_, p) in enumerate(pats)
                   for s in range(len(data) - len(p) + 1) if data[s:s + len(p)] == p)
    assert sorted(got) == brute


@pytest.mark.parametrize("seed", range(6))
def test_stream_scanner_standard_matches_scan_like_reference_property(seed):
    # core/tests/property/proptest_scan.rs:257-310: stream totals equal scan()
    # on the concatenation (compared by (start, pattern id)).
    r = np.random.default_rng(100 + seed)
    pats = [(0, bytes(r.integers(0, 256, size=int(r.integers(1, 8)), dtype=np.uint8))) for _ in range(3)]
    chunks = [bytes(r.integers(0, 256, size=int(r.integers(0, 200)), dtype=np.uint8)) for _ in range(10)]
    # plant the patterns across chunk boundaries
    data = bytearray(b"".join(chunks))
    for k in range(0, len(data) - 8, 97):
        p = pats[k % len(pats)][1]
        data[k:k + len(p)] = p
    data = bytes(data)
    cuts = np.cumsum([len(c) for c in chunks]).tolist()
    o = Oracle(pats)
    sc = StreamScannerOracle(o, _keep(pats), mode=1)
    got, prev = [], 0
    for c in cuts:
        got += sc.feed(data[prev:c])
        prev = c
    got += sc.finish()
    whole = o.scan_list(data, 1)
    assert sorted((s, p) for p, s, e in got) == sorted((s, p) for p, s, e in whole)
