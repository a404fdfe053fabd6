"""Continuous-batching request loop on the GPU: every request's match list is
the same alone and in the mix (exact: the scan is integer work), admission
respects slots and the in-flight byte budget, and routing never changes
results."""
import numpy as np
import pytest

from warpstate_amd import GpuMatcher, PatternSet
from warpstate_amd import workloads as W
from warpstate_amd.device import DeviceBuffer
from warpstate_amd.router import NCFG, config_tuning, supported_mask
from warpstate_amd.serve import Job, RequestLoop, poisson_arrivals, summarize

pytestmark = pytest.mark.gpu


def _sets():
    lits = W.literal_set(64, seed=0)
    kmers = W.kmer_set(500, seed=9)
    mixed, _, specs = W.mixed_set(70, 30, seed=13)
    out = []
    for kind_pats in ([(0, p) for p in lits], [(0, p) for p in kmers], mixed):
        b = PatternSet.builder()
        for kind, p in kind_pats:
            (b.literal if kind == 0 else b.regex)(p)
        out.append(b.build())
    pieces = lits + [p for k, p in mixed if k == 0] + [s.sample(np.random.default_rng(1)) for s in specs]
    return out, pieces, kmers


@pytest.fixture(scope="module")
def world():
    sets, pieces, kmers = _sets()
    pr = W.printable_input(24 << 20, 21)
    W.plant(pr, pieces, 2048, 22)
    ac = W.kmer_input(24 << 20, kmers, seed=23, plant_gap=2048)
    bufs = {"p": DeviceBuffer.from_array(pr), "a": DeviceBuffer.from_array(ac)}
    r = np.random.default_rng(24)
    jobs = []
    for i in range(60):
        s = int(r.integers(0, 3))
        n = int(np.exp(r.uniform(np.log(300), np.log(6 << 20))))
        off = int(r.integers(0, (24 << 20) - n))
        buf = bufs["a" if s == 1 else "p"]
        jobs.append(Job(i, s, buf.ptr + off, n, 0.0))
    return sets, bufs, jobs


def _alone(sets, jobs):
    out = {}
    ms = [GpuMatcher(s) for s in sets]
    for j in jobs:
        out[j.rid] = ms[j.patterns].scan_device_array(j.ptr, j.nbytes)
    return out


def test_results_are_the_same_alone_and_in_the_mix(world):
    sets, _, jobs = world
    alone = _alone(sets, jobs)
    infos = [GpuMatcher(s).info() for s in sets]
    masks = [supported_mask(i) for i in infos]
    cycle = {"k": 0}

    def route(batch):
        # Rotate through every supported configuration: routing must not
        # change any result.
        out = []
        for j in batch:
            while True:
                c = cycle["k"] % NCFG
                cycle["k"] += 1
                if masks[j.patterns][c]:
                    break
            out.append((c, config_tuning(c)))
        return out

    arrivals = poisson_arrivals(len(jobs), 4000.0, 25)
    for j, a in zip(jobs, arrivals):
        j.arrival = float(a)
    loop = RequestLoop(sets, route, slots=4, budget_bytes=16 << 20, keep_matches=True)
    try:
        outs = loop.run(jobs)
    finally:
        loop.close()
    assert len(outs) == len(jobs)
    for o in outs:
        assert np.array_equal(o.matches, alone[o.rid]), o.rid
    assert sum(len(v) for v in alone.values()) > 1000
    # Requests really overlapped on the device.
    ev = sorted([(o.admitted, 1) for o in outs] + [(o.done, -1) for o in outs])
    live = peak = 0
    for _, d in ev:
        live += d
        peak = max(peak, live)
    assert peak >= 2
    s = summarize(outs, [j.nbytes for j in jobs])
    assert s["requests"] == len(jobs) and s["p99_ms"] >= s["p50_ms"] > 0


def test_admission_respects_the_byte_budget(world):
    sets, _, jobs = world
    budget = 8 << 20
    loop = RequestLoop(sets, lambda b: [(1, config_tuning(1)) for _ in b], slots=3, budget_bytes=budget)
    try:
        for j in jobs:
            j.arrival = 0.0
        outs = loop.run(jobs)
    finally:
        loop.close()
    size = {j.rid: j.nbytes for j in jobs}
    ev = sorted([(o.admitted, 1, o.rid) for o in outs] + [(o.done, -1, o.rid) for o in outs],
                key=lambda e: (e[0], e[1]))
    live, nbytes = 0, 0
    for _, d, rid in ev:
        live += d
        nbytes += d * size[rid]
        assert live <= 3
        assert nbytes <= budget or live == 1
