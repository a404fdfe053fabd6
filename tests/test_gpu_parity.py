"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bit-exact (pattern_id, start, end) lists on seeded inputs, for every kernel
variant, plus the reference's own known-answer cases run on the device.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from warpstate_amd import (GpuMatcher, PatternSet, ScanItem, StreamPipeline, Tuning, VARIANT_DFA_GLOBAL,
                           VARIANT_DFA_LDS, VARIANT_HASH, WarpstateError, scan_batch)
from warpstate_amd import workloads as W

pytestmark = pytest.mark.gpu


def build(pats, ci=False):
    b = PatternSet.builder().case_insensitive(ci)
    for kind, p in pats:
        (b.literal if kind == 0 else b.regex)(p)
    return b.build()


def as_list(arr):
    return [(int(p), int(s), int(e)) for p, s, e in zip(arr["pattern_id"], arr["start"], arr["end"])]


def assert_same(got, exp, what=""):
    """Compact mismatch report (pytest's full list diff is too slow for big lists)."""
    if got == exp:
        return
    sg, se = set(got), set(exp)
    first = next((i for i, (a, b) in enumerate(zip(got, exp)) if a != b), min(len(got), len(exp)))
    raise AssertionError(
        f"{what}: {len(got)} vs {len(exp)} matches; first difference at {first}: "
        f"got {got[first:first + 3]} expected {exp[first:first + 3]}; "
        f"only-gpu {sorted(sg - se)[:5]} only-oracle {sorted(se - sg)[:5]}")


def variants_for(m):
    v = m.info()["variants_supported"]
    return [x for x in (VARIANT_HASH, VARIANT_DFA_GLOBAL, VARIANT_DFA_LDS) if v & (1 << x)]


def tunings_for(m):
    """Every kernel variant; the hash filter at each lane width (16, 32 and
    64 positions per lane: the benchmarks' 16 MiB, 64 MiB-chunk and 1 GiB
    shapes) on the device-sized pipeline, and at its default on the checked
    one."""
    out = []
    for v in variants_for(m):
        if v == VARIANT_HASH:
            out += [Tuning(variant=v, bytes_per_lane=b) for b in (16, 32, 64)]
            out.append(Tuning(variant=v, checked=True))
        else:
            out.append(Tuning(variant=v))
    return out


def check_all_variants(pats, data, mode=0, ci=False):
    ps = build(pats, ci)
    m = GpuMatcher(ps)
    exp = Oracle(pats, ci).scan_list(data, mode)
    for t in tunings_for(m):
        print(f"tuning={t} n={len(data)} pats={len(pats)} mode={mode}", flush=True)
        got = as_list(m.scan_array(data, mode="standard" if mode else "overlapping", tuning=t))
        assert_same(got, exp, f"tuning {t}")
    return exp


def test_literal_config_small():
    pats, data = W.literal_workload(size=1 << 20, seed=0)
    exp = check_all_variants([(0, p) for p in pats], data.tobytes())
    assert len(exp) > 5


def test_kmer_small():
    pats = W.kmer_set(10_000, seed=1)
    data = W.kmer_input(2 << 20, pats, seed=1)
    exp = check_all_variants([(0, p) for p in pats], data.tobytes())
    assert len(exp) > 100


def test_regex_grammar_small():
    srcs, data, _ = W.regex_workload(size=1 << 20, n=200, seed=2)
    data = W.regex_input(1 << 20, W.regex_set(200, 2), seed=2, plant_gap=2048)
    exp = check_all_variants([(1, s) for s in srcs], data.tobytes())
    assert len(exp) > 100


def test_mixed_set_standard_and_overlapping():
    pats, _, specs = W.mixed_set(70, 30, seed=3)
    pieces = [p for _, p in pats[:70]] + [s.sample(np.random.default_rng(9)) for s in specs]
    data = W.printable_input(1 << 19, 7)
    W.plant(data, pieces, 512, 8)
    check_all_variants(pats, data.tobytes(), mode=0)
    check_all_variants(pats, data.tobytes(), mode=1)


@pytest.mark.parametrize("npat,lmax", [(4, 8), (64, 16), (6000, 14)])
def test_standard_mode_strategies(npat, lmax):
    # 4 short literals -> MultiMemchr; 64 -> LeftmostLongest; 6000 -> HashScanner.
    alpha = np.frombuffer(b"ab", dtype=np.uint8)
    lits = W.random_literals(npat, 2 if npat < 100 else 10, lmax, alpha, seed=npat)
    data = alpha[np.random.default_rng(npat).integers(0, 2, size=200_000)].tobytes()
    o = Oracle([(0, l) for l in lits])
    assert o.strategy == {4: 2, 64: 1, 6000: 3}[npat]
    check_all_variants([(0, l) for l in lits], data, mode=1)


def test_reference_kats_on_device():
    # core/tests/correctness_guarantees.rs:15-41 overlapping abc / bc
    ps = PatternSet.builder().literal("abc").literal("bc").build()
    assert GpuMatcher(ps).scan(b"abc") == [(0, 0, 3), (1, 1, 3)]
    # correctness_guarantees.rs:257-299 duplicates, both ids in overlapping mode; one in scan
    ps = PatternSet.builder().literal("DUPE").literal("DUPE").build()
    assert GpuMatcher(ps).scan(b"xxDUPEyy") == [(0, 2, 6), (1, 2, 6)]
    assert ps.scan(b"xxDUPEyy") == [(0, 2, 6)]
    # all-null pattern, 3 overlapping matches (correctness_guarantees.rs:~230)
    ps = PatternSet.builder().literal_bytes(b"\x00\x00\x00").build()
    assert [m.start for m in ps.scan_overlapping(b"\x00" * 5)] == [0, 1, 2]
    # core/src/cpu/mod.rs:351-364: "aa" in "aaaa" non-overlapping
    ps = PatternSet.builder().literal("aa").build()
    assert ps.scan(b"aaaa") == [(0, 0, 2), (0, 2, 4)]
    # tests/unit/harness_regex.rs:65-68 and :77-80
    ps = PatternSet.builder().regex("[0-9]+").build()
    assert [(m.start, m.end) for m in ps.scan(b"abc 123 def 456")] == [(4, 7), (12, 15)]
    ps = PatternSet.builder().regex("foo|bar").build()
    assert [(m.start, m.end) for m in ps.scan(b"foo baz bar")] == [(0, 3), (8, 11)]
    # harness_regex.rs:28-31 case-insensitive group
    ps = PatternSet.builder().regex("(?i:hello)").build()
    assert len(ps.scan(b"Hello hello HELLO hElLo")) == 4
    # correctness_guarantees.rs:213-221 raw byte regex
    ps = PatternSet.builder().regex(r"\xFF").build()
    assert [(m.start, m.end) for m in ps.scan(bytes([0, 0xFF, 0]))] == [(1, 2)]


def test_edges_empty_short_tail():
    ps = PatternSet.builder().literal("needle").regex("ne+dle[0-9]{2}").build()
    m = GpuMatcher(ps)
    assert m.scan(b"") == []
    assert m.scan(b"ne") == []
    assert m.scan(b"needle") == [(0, 0, 6)]
    data = b"x" * 4093 + b"needle77"  # crosses a 4 KiB work unit
    exp = Oracle([(0, "needle"), (1, "ne+dle[0-9]{2}")]).scan_list(data)
    assert m.scan(data) == exp


def test_case_insensitive_literals():
    pats = [(0, b"Needle"), (0, b"HAYstack"), (0, b"ab")]
    data = b"needle NEEDLE haystack HayStack AB ab aB xx" * 50
    check_all_variants(pats, data, mode=0, ci=True)
    check_all_variants(pats, data, mode=1, ci=True)


def test_stream_pipeline_equals_whole():
    specs = W.regex_set(50, seed=11)
    data = W.regex_input(1 << 20, specs, seed=11, plant_gap=1024).tobytes()
    pats = [(1, s.source) for s in specs] + [(0, b"abc"), (0, b"zz_")]
    ps = build(pats)
    m = GpuMatcher(ps)
    whole = m.scan(data)
    overlap = max(s.max_len for s in specs)
    for chunk in (4096, 65536, 300_001):
        got = StreamPipeline(m, overlap).scan(data, chunk_size=chunk)
        assert_same(got, whole, f"chunk {chunk}")
    assert whole == [tuple(x) for x in Oracle(pats).scan_list(data)]


@pytest.mark.parametrize("checked", [False, True])
def test_stream_chains_spanning_many_blocks(checked):
    # A frequent bounded regex gives each chunk thousands of chain segments
    # of one (regex) group, spread over many 256-lane blocks of the chain
    # walk; the state carried from the previous chunk must read the same in
    # every segment and in both passes of a lane (count, then emit).
    rng = np.random.default_rng(77)
    data = bytearray(W.printable_input(1 << 20, 78).tobytes())
    for pos in range(0, len(data) - 8, 11):
        data[pos:pos + 2] = b"q" + bytes([0x30 + int(rng.integers(0, 10))])
    data = bytes(data)
    pats = [(1, b"q[0-9][a-z0-9]{0,3}"), (1, b"[0-9]q[0-9]"), (0, b"q1")]
    m = GpuMatcher(build(pats))
    exp = [tuple(x) for x in Oracle(pats).scan_list(data)]
    assert len(exp) > 100_000
    for chunk in (65_536, 100_003):
        got = StreamPipeline(m, 16).scan_array(data, chunk_size=chunk, tuning=Tuning(checked=checked))
        assert_same(as_list(got), exp, f"chunk {chunk} checked={checked}")


@pytest.mark.parametrize("npat", [4, 64, 6000])
def test_stream_standard_mode_equals_whole(npat):
    # PatternSet::scan semantics through StreamPipeline over many chunks: the
    # literal reductions (MultiMemchr, LeftmostLongest, HashScanner) span the
    # chunk edges, so the result is the whole input's.
    alpha = np.frombuffer(b"ab", dtype=np.uint8)
    lmin, lmax = {4: (2, 12), 64: (4, 12), 6000: (12, 14)}[npat]
    lits = W.random_literals(npat, lmin, lmax, alpha, seed=npat + 1)
    data = alpha[np.random.default_rng(npat + 1).integers(0, 2, size=150_000)].tobytes()
    pats = [(0, l) for l in lits] + ([(1, b"ab{3,5}a")] if npat == 64 else [])
    ps = build(pats)
    m = GpuMatcher(ps)
    exp = [tuple(x) for x in Oracle(pats).scan_list(data, mode=1)]
    overlap = max(len(l) for l in lits) + 8
    for chunk in (4096, 65_537):
        got = StreamPipeline(m, overlap).scan_array(data, mode="standard", chunk_size=chunk)
        assert_same(as_list(got), exp, f"chunk {chunk}")
    with pytest.raises(WarpstateError):
        StreamPipeline(m, 0).scan_array(data, mode="standard", chunk_size=4096)


def test_stream_boundary_kat():
    # tests/unit/mod.rs:11-54: "legendary" over a 1024-byte chunk edge
    ps = PatternSet.builder().regex("legendary").build()
    data = bytearray(b"." * 2048)
    data[1020:1029] = b"legendary"
    got = StreamPipeline(GpuMatcher(ps), 16).scan(bytes(data), chunk_size=1024)
    assert got == [(0, 1020, 1029)]


def test_batch_per_item_exact():
    pats, _, specs = W.mixed_set(70, 30, seed=3)
    rng = np.random.default_rng(5)
    pieces = [p for _, p in pats[:70]] + [s.sample(rng) for s in specs]
    ids, offs, sizes, buf = W.batch_items(500, seed=4, pieces=pieces, plant_gap=256)
    items = [ScanItem(int(i), buf[o:o + s].tobytes()) for i, o, s in zip(ids, offs, sizes)]
    m = GpuMatcher(build(pats))
    got = scan_batch(m, items)
    o = Oracle(pats)
    exp = []
    for it in items:
        exp += [(it.id, x) for x in o.scan_list(it.data)]
    assert_same([(t.source_id, tuple(t.matched)) for t in got], exp, "batch")
    # tests/unit/gpu_batch.rs:6-65: a match split across two items is not reported
    m2 = GpuMatcher(PatternSet.builder().literal("needle").build())
    got2 = scan_batch(m2, [ScanItem(1, b"xxnee"), ScanItem(2, b"dlexx"), ScanItem(3, b"needle")])
    assert [(t.source_id, tuple(t.matched)) for t in got2] == [(3, (0, 0, 6))]


def test_batch_device_items_unaligned():
    # Items back to back in one device allocation (arbitrary alignment),
    # ids out of order, some empty: TaggedMatch order is (source_id, start,
    # pattern_id, end) and every item is scanned on its own.
    from warpstate_amd import scan_batch_device
    from warpstate_amd.device import DeviceBuffer

    pats, _, specs = W.mixed_set(70, 30, seed=8)
    rng = np.random.default_rng(9)
    pieces = [p for _, p in pats[:70]] + [s.sample(rng) for s in specs]
    ids, offs, sizes, buf = W.batch_items(700, seed=10, pieces=pieces, plant_gap=300)
    sizes = sizes.copy()
    sizes[::37] = 0
    ids = rng.permutation(np.arange(700, dtype=np.uint64) * 3 + 11)
    dev = DeviceBuffer.from_array(buf)
    m = GpuMatcher(build(pats))
    got = scan_batch_device(m, ids, (dev.ptr + offs).astype(np.uint64), sizes)
    o = Oracle(pats)
    exp = []
    for i in np.argsort(ids, kind="stable"):
        chunk = buf[offs[i]:offs[i] + sizes[i]].tobytes()
        exp += [(int(ids[i]),) + x for x in o.scan_list(chunk)]
    got_l = [(int(a), int(b), int(c), int(d)) for a, b, c, d in
             zip(got["source_id"], got["pattern_id"], got["start"], got["end"])]
    assert len(exp) > 200
    assert_same(got_l, exp, "device batch")


def test_stream_scan_device_input_equals_whole():
    from warpstate_amd.device import DeviceBuffer

    srcs, data, maxlen = W.regex_workload(size=3 << 20, n=40, seed=14)
    m = GpuMatcher(build([(1, s.encode()) for s in srcs]))
    whole = m.scan_array(data.tobytes())
    dev = DeviceBuffer.from_array(data)
    out = np.zeros(len(whole) + 16, dtype=whole.dtype)
    n = StreamPipeline(m, maxlen).scan_device(dev.ptr, len(data), out.ctypes.data, len(out),
                                              chunk_size=256 << 10, out_on_device=False)
    assert n == len(whole) and len(whole) > 50
    assert np.array_equal(out[:n], whole)


def test_ordering_many_matches_spread_and_clustered():
    # Many matches spread over the input (the bucketed ordering) and the same
    # count packed into a few KiB of a large input (buckets overflow; the
    # ordering falls back to the radix sort): both must equal the oracle.
    lits = W.random_literals(300, 4, 9, W.LIT_ALPHABET, seed=31)
    data = W.printable_input(4 << 20, 32)
    W.plant(data, lits, 48, 33)
    m = GpuMatcher(build([(0, l) for l in lits]))
    got = m.scan_array(data.tobytes())
    exp = Oracle([(0, l) for l in lits]).scan(data, 0)
    assert len(exp) > 50_000
    assert np.array_equal(got, exp)

    big = np.zeros(64 << 20, dtype=np.uint8)
    big[:10_000] = np.frombuffer(b"ab" * 5000, dtype=np.uint8)
    big[(40 << 20):(40 << 20) + 6] = np.frombuffer(b"xbabax", dtype=np.uint8)
    m2 = GpuMatcher(build([(0, b"ab"), (0, b"ba")]))
    got2 = as_list(m2.scan_array(big.tobytes()))
    exp2 = sorted([(0, i, i + 2) for i in range(0, 10_000, 2)] + [(1, i, i + 2) for i in range(1, 9_999, 2)] +
                  [(1, (40 << 20) + 1, (40 << 20) + 3), (0, (40 << 20) + 2, (40 << 20) + 4),
                   (1, (40 << 20) + 3, (40 << 20) + 5)],
                  key=lambda t: (t[1], t[0], t[2]))
    assert_same(got2, exp2, "clustered")


def test_device_sized_and_checked_pipelines_agree():
    # Overlapping scans run device-sized by default (one synchronisation per
    # call); the host-checked pipeline must give the same lists, whole and
    # chunked, and both must equal the oracle.
    pats, _, specs = W.mixed_set(150, 60, seed=41)
    rng = np.random.default_rng(42)
    pieces = [p for k, p in pats if k == 0] + [s.sample(rng) for s in specs]
    data = W.printable_input(3 << 20, 43)
    W.plant(data, pieces, 700, 44)
    m = GpuMatcher(build(pats))
    dev = m.scan_array(data.tobytes())
    chk = m.scan_array(data.tobytes(), tuning=Tuning(checked=True))
    exp = Oracle(pats).scan(data, 0)
    assert len(exp) > 3000
    assert np.array_equal(dev, exp) and np.array_equal(chk, exp)
    maxlen = max(s.max_len for s in specs) + 24
    sp = StreamPipeline(m, maxlen)
    sd = sp.scan_array(data.tobytes(), chunk_size=300_000)
    sc = sp.scan_array(data.tobytes(), chunk_size=300_000, tuning=Tuning(checked=True))
    assert np.array_equal(sd, exp) and np.array_equal(sc, exp)


def test_device_sized_pipeline_redoes_on_overflow():
    # 500,000 regex candidates overflow the device-sized capacity for a
    # 1 MB input; the scan is redone on the checked pipeline, same answer.
    data = b"ab" * 500_000 + b"ab7"
    m = GpuMatcher(build([(1, b"ab[0-9]")]))
    got = as_list(m.scan_array(data))
    assert got == [(0, 1_000_000, 1_000_003)]


def test_graph_replays_match_eager():
    # A repeated device scan is captured into a hipGraph on its second call
    # and replayed afterwards; every replay must equal the eager result, with
    # other eager work and synchronisations between the replays.
    from warpstate_amd import AutoMatcherConfig, Context
    from warpstate_amd.device import DeviceBuffer

    pats, _, specs = W.mixed_set(120, 40, seed=51)
    rng = np.random.default_rng(52)
    pieces = [p for k, p in pats if k == 0] + [s.sample(rng) for s in specs]
    data = W.printable_input(2 << 20, 53)
    W.plant(data, pieces, 500, 54)
    other = W.printable_input(1 << 20, 55)
    W.plant(other, pieces, 300, 56)
    ctx = Context(0, AutoMatcherConfig())
    m = GpuMatcher(build(pats), context=ctx)
    exp = Oracle(pats).scan(data, 0)
    exp_other = Oracle(pats).scan(other, 0)
    dev, dev_other = DeviceBuffer.from_array(data), DeviceBuffer.from_array(other)
    for k in range(5):
        got = m.scan_device_array(dev.ptr, len(data))
        assert np.array_equal(got, exp), k
        assert np.array_equal(m.scan_array(other.tobytes()), exp_other), k
    assert ctx.graph_count >= 1
    # Stream steady state: chunked scans of device memory replay too.
    maxlen = max(s.max_len for s in specs) + 24
    sp = StreamPipeline(m, maxlen)
    out = DeviceBuffer(12 * (len(exp) + 64))
    for k in range(4):
        n = sp.scan_device(dev.ptr, len(data), out.ptr, len(exp) + 64, chunk_size=400_000)
        assert n == len(exp)
        got = out.download(np.uint8, 12 * n).view(exp.dtype)
        assert np.array_equal(got, exp), k
        assert np.array_equal(m.scan_device_array(dev_other.ptr, len(other)), exp_other), k
    assert ctx.graph_count >= 3


def test_overflow_is_an_error_not_a_fallback():
    from warpstate_amd import AutoMatcherConfig, Context
    ctx = Context(0, AutoMatcherConfig(max_matches=100))
    ps = PatternSet.builder().literal("a").build()
    with pytest.raises(WarpstateError) as e:
        GpuMatcher(ps, context=ctx).scan(b"a" * 1000)
    assert e.value.variant == "MatchBufferOverflow"


@pytest.mark.parametrize("chunk", [1 << 20, 96 << 10])
def test_literal_stream_runs_and_replays(chunk):
    # A literal-only set over device memory in chunks: 4 runs (segments
    # passed to the init kernel by value) or 43 runs (more than it takes:
    # segments copied from pinned staging), each ordered through the
    # fixed-capacity buckets.
    # Eager, captured and replayed calls all equal the oracle's whole scan.
    from warpstate_amd import AutoMatcherConfig, Context
    from warpstate_amd.device import DeviceBuffer

    lits = W.random_literals(400, 5, 14, W.LIT_ALPHABET, seed=61)
    data = W.printable_input(4 << 20, 62)
    W.plant(data, lits, 300, 63)
    exp = Oracle([(0, l) for l in lits]).scan(data, 0)
    assert len(exp) > 10_000
    ctx = Context(0, AutoMatcherConfig())
    m = GpuMatcher(build([(0, l) for l in lits]), context=ctx)
    sp = StreamPipeline(m, max(len(l) for l in lits))
    dev = DeviceBuffer.from_array(data)
    out = DeviceBuffer(12 * (len(exp) + 64))
    exp_short = Oracle([(0, l) for l in lits]).scan(data[: len(data) - 7], 0)
    for k in range(4):
        n = sp.scan_device(dev.ptr, len(data), out.ptr, len(exp) + 64, chunk_size=chunk)
        assert n == len(exp), k
        got = out.download(np.uint8, 12 * n).view(exp.dtype)
        assert np.array_equal(got, exp), k
        # a whole scan in between (another graph, other segments)
        assert np.array_equal(m.scan_device_array(dev.ptr, len(data) - 7), exp_short), k
    assert ctx.graph_count >= 2


def test_state_block_with_profiling_toggled_and_many_runs():
    # The call's state block (flags, total, one pair of kernel stamps per
    # run) grows with the run count when profiling is on; calls alternate
    # profiling off and on, eager and replayed, with 1 and 64 runs. Every
    # result equals the oracle and profiled calls report the filter time.
    from warpstate_amd import AutoMatcherConfig, Context
    from warpstate_amd.device import DeviceBuffer

    pats, _, specs = W.mixed_set(80, 20, seed=71)
    rng = np.random.default_rng(72)
    pieces = [p for k, p in pats if k == 0] + [s.sample(rng) for s in specs]
    data = W.printable_input(3 << 20, 73)
    W.plant(data, pieces, 400, 74)
    exp = Oracle(pats).scan(data, 0)
    ctx = Context(0, AutoMatcherConfig())
    m = GpuMatcher(build(pats), context=ctx)
    sp = StreamPipeline(m, max(s.max_len for s in specs) + 24)
    dev = DeviceBuffer.from_array(data)
    out = DeviceBuffer(12 * (len(exp) + 64))
    for k in range(8):
        prof = k % 2 == 1
        ctx.set_profiling(prof)
        got = m.scan_device_array(dev.ptr, len(data))
        assert np.array_equal(got, exp), k
        if prof:
            assert ctx.last_timings().prefilter_ms > 0, k
        n = sp.scan_device(dev.ptr, len(data), out.ptr, len(exp) + 64, chunk_size=48 << 10)
        assert n == len(exp), k
        assert np.array_equal(out.download(np.uint8, 12 * n).view(exp.dtype), exp), k
        if prof:
            assert ctx.last_timings().prefilter_ms > 0, k
    ctx.set_profiling(False)
    assert ctx.graph_count >= 4


def test_filter_hit_count_is_summed_over_every_block():
    # The filter sums its hits per block and adds each block's sum once to
    # the call's counter. Every match start is a hit, the count is the same
    # on the device-sized and checked pipelines, eager and replayed, and a
    # scan of twice the input (the same bytes again) counts exactly twice
    # as many hits.
    lits, data = W.literal_workload(size=4 << 20, seed=3)
    m = GpuMatcher(build([(0, p) for p in lits]))
    ctx = m.ctx
    ctx.set_profiling(True)
    try:
        counts = []
        for tuning in (None, None, None, Tuning(checked=True)):
            got = m.scan_array(data.tobytes(), tuning=tuning)
            counts.append(int(ctx.last_timings().hits))
        assert len(set(counts)) == 1, counts
        starts = len(set(int(s) for s in got["start"]))
        assert starts > 20 and counts[0] >= starts
        # No literal spans the seam (they are 4-16 bytes over [a-z_], the
        # input is printable with planted copies), so hits double exactly
        # when every window of the doubled input is hashed; the seam's
        # windows are checked by the match count instead.
        twice = np.concatenate([data, data])
        got2 = m.scan_array(twice.tobytes())
        exp2 = Oracle([(0, p) for p in lits]).scan(twice, 0)
        assert np.array_equal(got2, exp2)
        h2 = int(ctx.last_timings().hits)
        assert 2 * counts[0] <= h2 <= 2 * counts[0] + 16
    finally:
        ctx.set_profiling(False)


def test_repeat_calls_skip_init_only_when_state_is_clean():
    # A device-sized call whose segments equal the previous call's skips
    # the init kernel (the previous close kernel reset the state block).
    # Checked, batch and stream calls in between reuse the counters and the
    # segment list, so the next call must initialise again. Every result
    # equals the oracle's, hits included.
    lits, data = W.literal_workload(size=2 << 20, seed=8)
    pats = [(0, p) for p in lits]
    m = GpuMatcher(build(pats))
    raw = data.tobytes()
    exp = Oracle(pats).scan(data, 0)
    half = Oracle(pats).scan(data[: len(raw) // 2], 0)
    ctx = m.ctx
    ctx.set_profiling(True)
    try:
        hits = set()
        for step in ["dev", "dev", "dev", "checked", "dev", "batch", "dev", "stream", "dev", "half", "dev", "dev"]:
            if step == "checked":
                assert np.array_equal(m.scan_array(raw, tuning=Tuning(checked=True)), exp)
            elif step == "batch":
                got = scan_batch(m, [ScanItem(5, raw)])
                assert len(got) > 20 and all(t.source_id == 5 for t in got)
            elif step == "stream":
                assert np.array_equal(StreamPipeline(m, 32).scan_array(raw, chunk_size=700_000), exp)
            elif step == "half":
                assert np.array_equal(m.scan_array(raw[: len(raw) // 2]), half)
            else:
                assert np.array_equal(m.scan_array(raw), exp)
                hits.add(int(ctx.last_timings().hits))
        assert len(hits) == 1
    finally:
        ctx.set_profiling(False)


@pytest.mark.parametrize("kind", ["mixed", "literals", "look"])
def test_batch_plans_device_sized_replayed_and_device_output(kind):
    # A batch plan scanned again and again (device-sized, then captured and
    # replayed), into host and device memory, equals the unplanned batch and
    # every item's own oracle scan; ids out of order and empty items.
    from warpstate_amd import AutoMatcherConfig, BatchPlan, Context, scan_batch_device, scan_batch_planned
    from warpstate_amd.device import DeviceBuffer

    if kind == "mixed":
        pats, _, specs = W.mixed_set(300, 100, seed=81)
    elif kind == "literals":
        pats, specs = [(0, l) for l in W.random_literals(500, 4, 20, W.LIT_ALPHABET, seed=82)], []
    else:
        pats, specs = [(1, r"^ab"), (1, r"b\b"), (1, r"^x*"), (0, b"abc")], []
    rng = np.random.default_rng(83)
    pieces = [p for k, p in pats if k == 0] + [s.sample(rng) for s in specs] + [b"ab", b"xab"]
    ids, offs, sizes, buf = W.batch_items(2000, seed=84, pieces=pieces, plant_gap=500)
    sizes = sizes.copy()
    sizes[::53] = 0
    ids = rng.permutation(np.arange(len(ids), dtype=np.uint64) * 5 + 3)
    dev = DeviceBuffer.from_array(buf)
    ptrs = (dev.ptr + offs).astype(np.uint64)
    ctx = Context(0, AutoMatcherConfig())
    m = GpuMatcher(build(pats), context=ctx)
    o = Oracle(pats)
    exp_l = []
    for i in np.argsort(ids, kind="stable"):
        chunk = buf[offs[i]:offs[i] + sizes[i]]
        exp_l += [(int(ids[i]), int(p), int(s), int(e)) for p, s, e in
                  zip(*[o.scan(chunk, 0)[f] for f in ("pattern_id", "start", "end")])]
    assert len(exp_l) > 1000

    def as_l(a):
        return [(int(x), int(y), int(z), int(w)) for x, y, z, w in
                zip(a["source_id"], a["pattern_id"], a["start"], a["end"])]

    assert as_l(scan_batch_device(m, ids, ptrs, sizes)) == exp_l
    plan = BatchPlan(ctx, ids, ptrs, sizes)
    out = DeviceBuffer(24 * (len(exp_l) + 64))
    for k in range(4):
        assert as_l(scan_batch_planned(m, plan)) == exp_l, k
        n = scan_batch_planned(m, plan, out_ptr=out.ptr, capacity=len(exp_l) + 64)
        assert n == len(exp_l), k
        assert as_l(out.download(np.uint8, 24 * n).view(N_TAGGED())) == exp_l, k
    assert ctx.graph_count >= 2
    # the checked pipeline over the same plan
    assert as_l(scan_batch_planned(m, plan, tuning=Tuning(checked=True))) == exp_l
    plan.close()


def N_TAGGED():
    from warpstate_amd import _native as N

    return N.TAGGED_DTYPE
