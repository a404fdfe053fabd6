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


def test_served_requests_equal_the_oracle_on_every_request_set():
    # The request loop's own workload (core/src/router.rs:205-225 serves one
    # AutoMatcher::scan per request): the 20 request pattern sets (10 to
    # 10,000 patterns, literal fractions 1 to 0, ACGT k-mer sets among them),
    # log-uniform request sizes over planted device-resident pools, Poisson
    # arrivals, a random supported configuration per request. Every served
    # list whose oracle work (bytes x (regexes + 1)) is under 2^28 is
    # compared exactly with the oracle, and a seeded sample of 5 above it
    # (up to 2^33).
    from oracle.oracle import Oracle
    from warpstate_amd.router import DevicePools, build_patternset, make_requests, request_pattern_sets

    specs = request_pattern_sets()
    sets = [build_patternset(sp) for sp in specs]
    pools = DevicePools(specs, size=32 << 20, seed=31)
    reqs = make_requests(300, specs, pools, seed=32)
    infos = [GpuMatcher(s).info() for s in sets]
    masks = [supported_mask(i) for i in infos]
    rng = np.random.default_rng(33)

    def route(batch):
        out = []
        for j in batch:
            c = int(rng.choice(np.nonzero(masks[j.patterns])[0]))
            out.append((c, config_tuning(c)))
        return out

    arrivals = poisson_arrivals(len(reqs), 3000.0, 34)
    jobs = [Job(q.rid, q.set_index, pools.ptr(specs[q.set_index].kind, q.density, q.offset), q.nbytes, float(a),
                density=pools.density[(specs[q.set_index].kind, q.density)]) for q, a in zip(reqs, arrivals)]
    loop = RequestLoop(sets, route, slots=4, budget_bytes=64 << 20, keep_matches=True)
    try:
        outs = {o.rid: o for o in loop.run(jobs)}
    finally:
        loop.close()
    assert len(outs) == len(jobs)

    work = {q.rid: q.nbytes * (infos[q.set_index]["regex_count"] + 1) for q in reqs}
    under = [q for q in reqs if work[q.rid] <= (1 << 28)]
    over = [q for q in reqs if (1 << 28) < work[q.rid] <= (1 << 33)]
    sample = [over[int(i)] for i in np.random.default_rng(35).choice(len(over), size=min(5, len(over)),
                                                                      replace=False)]
    assert len(under) > 150 and len({q.set_index for q in under}) == len(specs)
    host = {key: buf.download(np.uint8, pools.size) for key, buf in pools.dev.items()}
    oracles = {}
    total = 0
    for q in under + sample:
        kind = specs[q.set_index].kind
        data = host[(kind, q.density)][q.offset:q.offset + q.nbytes]
        if q.set_index not in oracles:
            oracles[q.set_index] = Oracle(specs[q.set_index].patterns)
        exp = oracles[q.set_index].scan_mt(data, 0, 16)
        got = outs[q.rid].matches
        assert np.array_equal(got, exp), (q.rid, q.set_index, q.nbytes, outs[q.rid].config, len(got), len(exp))
        total += len(exp)
    assert total > 10_000
