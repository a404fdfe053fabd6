"""GPU parity at the benchmarked shapes: the configurations the benchmark
measures, at (or near) their full sizes, against the threaded CPU oracle.

* 16 MiB literal scan, 64 literals (the default 16 positions per lane);
* 1 GiB k-mer scan, 10,000 literals, default tuning: the 64-positions-per-lane
  filter that the headline measures, with its exact match list;
* 200-regex grammar set streamed in 64 MiB chunks (overlap = longest match)
  over 192 MiB of device memory, eager and replayed;
* batch of log-normal items against the 1,000-pattern mixed set (700
  literals + 300 grammar regexes), every item against its own oracle scan.

Match lists are compared exactly (integer work).
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from warpstate_amd import GpuMatcher, PatternSet, StreamPipeline, scan_batch_device
from warpstate_amd import workloads as W
from warpstate_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu

THREADS = 16


def build(pats):
    b = PatternSet.builder()
    for kind, p in pats:
        (b.literal if kind == 0 else b.regex)(p)
    return b.build()


def first_diff(got, exp):
    n = min(len(got), len(exp))
    bad = np.nonzero(got[:n] != exp[:n])[0]
    i = int(bad[0]) if len(bad) else n
    return f"{len(got)} vs {len(exp)} matches; first difference at {i}: got {got[i:i + 2]} expected {exp[i:i + 2]}"


def test_literal_16mib_default_tuning():
    pats, data = W.literal_workload(size=16 << 20, seed=0)
    pp = [(0, p) for p in pats]
    exp = Oracle(pp).scan_mt(data, 0, THREADS)
    dev = DeviceBuffer.from_array(data)
    got = GpuMatcher(build(pp)).scan_device_array(dev.ptr, len(data))
    print(f"literal 16 MiB: {len(exp)} matches", flush=True)
    assert len(exp) > 200
    assert np.array_equal(got, exp), first_diff(got, exp)


def test_kmer_1gib_default_tuning():
    # The headline workload: 1 GiB, 10,000 k-mers, 64 positions per lane.
    pats, data = W.kmer_workload(1 << 30)
    pp = [(0, p) for p in pats]
    print("k-mer 1 GiB: oracle", flush=True)
    exp = Oracle(pp).scan_mt(data, 0, THREADS)
    dev = DeviceBuffer.from_array(data)
    del data
    m = GpuMatcher(build(pp))
    print("k-mer 1 GiB: gpu", flush=True)
    for k in range(3):  # eager, captured, replayed
        got = m.scan_device_array(dev.ptr, 1 << 30)
        assert np.array_equal(got, exp), f"call {k}: " + first_diff(got, exp)
    assert len(exp) == 297_412


def test_regex_stream_64mib_chunks():
    # The regex workload's shape: 200 grammar regexes, StreamPipeline with
    # 64 MiB chunks and overlap = the longest match, over 192 MiB.
    size = 192 << 20
    srcs, data, maxlen = W.regex_workload(size=size, n=200, seed=2)
    pp = [(1, s) for s in srcs]
    print("regex 192 MiB: oracle", flush=True)
    exp = Oracle(pp).scan_mt(data, 0, THREADS)
    assert len(exp) > 5000
    dev = DeviceBuffer.from_array(data)
    m = GpuMatcher(build(pp))
    sp = StreamPipeline(m, maxlen)
    out = DeviceBuffer(12 * (len(exp) + 1024))
    print("regex 192 MiB: gpu", flush=True)
    for k in range(3):
        n = sp.scan_device(dev.ptr, size, out.ptr, len(exp) + 1024, chunk_size=64 << 20)
        got = out.download(np.uint8, 12 * n).view(exp.dtype)
        assert np.array_equal(got, exp), f"call {k}: " + first_diff(got, exp)


def test_batch_1000_pattern_set():
    # The batch workload's pattern set (700 literals 4-24 bytes + 300 grammar
    # regexes) over 3,000 log-normal items in one device allocation.
    pats, _, specs = W.mixed_set(700, 300, seed=3)
    rng = np.random.default_rng(4)
    pieces = [p for k, p in pats if k == 0] + [s.sample(rng) for s in specs]
    ids, offs, sizes, buf = W.batch_items(3000, seed=4, pieces=pieces, plant_gap=2048)
    dev = DeviceBuffer.from_array(buf)
    m = GpuMatcher(build(pats))
    got = scan_batch_device(m, ids, (dev.ptr + offs).astype(np.uint64), sizes)
    o = Oracle(pats)
    exp = []
    for i in range(len(ids)):
        chunk = buf[offs[i]:offs[i] + sizes[i]]
        a = o.scan_mt(chunk, 0, THREADS)
        exp += [(int(ids[i]), int(p), int(s), int(e)) for p, s, e in zip(a["pattern_id"], a["start"], a["end"])]
    got_l = [(int(a), int(b), int(c), int(d)) for a, b, c, d in
             zip(got["source_id"], got["pattern_id"], got["start"], got["end"])]
    print(f"batch 3,000 items: {len(exp)} matches", flush=True)
    assert len(exp) > 5000
    assert got_l == exp
