"""StreamScanner (core/src/stream.rs) on the device against the oracle's
restatement: the reference's own known-answer cases, then seeded random
chunkings of literal and mixed sets in both modes, host and device chunks."""
import json
import os

import numpy as np
import pytest

from oracle.oracle import Oracle, StreamScannerOracle
from warpstate_amd import GpuMatcher, PatternSet, StreamScanner
from warpstate_amd import workloads as W
from warpstate_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def build(pats):
    b = PatternSet.builder()
    for kind, p in pats:
        (b.literal_bytes if kind == 0 else b.regex)(p)
    return b.build()


def keep_of(pats):
    return max((len(p) for k, p in pats if k == 0), default=0)


@pytest.mark.parametrize("kat", json.load(open(os.path.join(HERE, "golden", "stream_kats.json"))),
                         ids=lambda k: k["name"])
def test_stream_scanner_reference_kats(kat):
    pats = [(k, bytes.fromhex(h)) for k, h in kat["patterns"]]
    sc = StreamScanner(GpuMatcher(build(pats)))
    for feed, exp in zip(kat["feeds"], kat["expected"]):
        assert sc.feed(bytes.fromhex(feed)) == [tuple(e) for e in exp], kat["source"]
    sc.close()


def feeds_of(data, rng, n):
    cuts = sorted(set(rng.integers(0, len(data), size=n).tolist()))
    out, prev = [], 0
    for c in cuts + [len(data)]:
        out.append(data[prev:c])
        prev = c
    return out


@pytest.mark.parametrize("mode", ["standard", "overlapping"])
@pytest.mark.parametrize("kind", ["literals", "mixed"])
def test_stream_scanner_random_chunkings(mode, kind):
    rng = np.random.default_rng(7)
    if kind == "literals":
        lits, data = W.literal_workload(size=1 << 18, seed=3)
        pats = [(0, p) for p in lits]
        data = data.tobytes()
    else:
        pats, _, specs = W.mixed_set(60, 20, seed=5)
        data = W.regex_input(1 << 18, specs, seed=5, plant_gap=512).tobytes()
    m = GpuMatcher(build(pats))
    code = 1 if mode == "standard" else 0
    for trial, n in enumerate((3, 40, 400)):
        chunks = feeds_of(data, rng, n)
        print(f"kind={kind} mode={mode} feeds={len(chunks)}", flush=True)
        exp = StreamScannerOracle(Oracle(pats), keep_of(pats), mode=code)
        sc = StreamScanner(m, mode=mode)
        for i, ch in enumerate(chunks):
            assert sc.feed(ch) == exp.feed(ch), f"feed {i} of {len(chunks)}"
        assert sc.finish() == exp.finish()
        assert sc.processed_bytes == len(data)
        sc.close()


def test_stream_scanner_device_chunks_equal_host_chunks():
    lits, data = W.literal_workload(size=1 << 18, seed=9)
    pats = [(0, p) for p in lits]
    m = GpuMatcher(build(pats))
    buf = DeviceBuffer.from_array(data)
    rng = np.random.default_rng(1)
    cuts = sorted(set(rng.integers(0, len(data), size=30).tolist())) + [len(data)]
    host, dev = StreamScanner(m), StreamScanner(m)
    prev = 0
    raw = data.tobytes()
    for c in cuts:
        assert dev.feed_device(buf.ptr + prev, c - prev) == host.feed(raw[prev:c])
        prev = c
    assert dev.finish() == host.finish()
    host.close()
    dev.close()
    buf.free()
