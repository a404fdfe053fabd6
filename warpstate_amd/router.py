"""Kernel routing for the scan path: exhaustive autotuner, learned MLP router,
AutoMatcher, and the continuous-batching request loop.

The reference routes every request by one size threshold
(AutoMatcher::scan, core/src/router.rs:205-225; threshold default
core/src/config.rs:4-35). Here each request picks one of WS_ROUTER_CONFIGS
kernel configurations (variant x lane width x occupancy x pipeline). The exhaustive
autotuner (ws_autotune) times all of them; the learned router is a small FP32
MLP trained on those timings by the library's own HIP kernels
(ws_router_train) and predicts the fastest configuration from 7 request
features. Routing never changes results: every configuration returns the same
match list.
"""
from __future__ import annotations

import ctypes
import queue
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N
from .api import AutoMatcherConfig, Context, GpuMatcher, PatternSet, Tuning, default_context
from . import workloads as W

NCFG = 16
NFEAT = 7


def config_tuning(i: int) -> Tuning:
    t = N.WsTuning()
    N.check(N.load().ws_config_tuning(i, ctypes.byref(t)))
    return Tuning(t.variant, t.bytes_per_lane, t.blocks_per_cu, bool(t.flags & N.TUNING_CHECKED))


def features(info: dict, nbytes: int, matches_per_mib: float) -> np.ndarray:
    ci = N.WsPatternsInfo(**info)
    out = (ctypes.c_float * 8)()
    N.check(N.load().ws_router_features(ctypes.byref(ci), int(nbytes), float(matches_per_mib), out))
    return np.array(list(out), dtype=np.float32)


def supported_mask(info: dict) -> np.ndarray:
    m = np.zeros(NCFG, dtype=np.float32)
    for i in range(NCFG):
        if info["variants_supported"] & (1 << config_tuning(i).variant):
            m[i] = 1.0
    return m


class Router:
    """The learned cost model (ws_router_*): features -> predicted ms per config."""

    def __init__(self, device: int = 0, seed: int = 0):
        h = ctypes.c_void_p()
        N.check_router(N.load().ws_router_create(device, seed, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            N.load().ws_router_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def train(self, x: np.ndarray, log_ms: np.ndarray, mask: np.ndarray, epochs: int = 400, batch: int = 256,
              lr: float = 2e-3, seed: int = 0) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(log_ms, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.float32)
        loss = ctypes.c_float(0)
        N.check_router(N.load().ws_router_train(self.h, x.ctypes.data, y.ctypes.data, m.ctypes.data, len(x), epochs,
                                                batch, lr, seed, ctypes.byref(loss)))
        return float(loss.value)

    def predict(self, x: np.ndarray, mask: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        m = np.ascontiguousarray(np.atleast_2d(mask), dtype=np.float32)
        pred = np.zeros((len(x), NCFG), dtype=np.float32)
        choice = np.zeros(len(x), dtype=np.uint32)
        N.check_router(N.load().ws_router_predict(self.h, x.ctypes.data, len(x), m.ctypes.data, pred.ctypes.data,
                                                  choice.ctypes.data))
        return choice, pred

    def params(self) -> np.ndarray:
        p = np.zeros(N.load().ws_router_param_count(), dtype=np.float32)
        N.check_router(N.load().ws_router_get_params(self.h, p.ctypes.data, len(p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        N.check_router(N.load().ws_router_set_params(self.h, p.ctypes.data, len(p)))

    def forward(self, xn: np.ndarray) -> np.ndarray:
        xn = np.ascontiguousarray(xn, dtype=np.float32)
        out = np.zeros((len(xn), NCFG), dtype=np.float32)
        N.check_router(N.load().ws_router_forward(self.h, xn.ctypes.data, len(xn), out.ctypes.data))
        return out

    def gradient(self, xn, y, m, scale: float) -> np.ndarray:
        xn = np.ascontiguousarray(xn, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        m = np.ascontiguousarray(m, dtype=np.float32)
        g = np.zeros(N.load().ws_router_param_count(), dtype=np.float32)
        N.check_router(N.load().ws_router_gradient(self.h, xn.ctypes.data, y.ctypes.data, m.ctypes.data, len(xn),
                                                   scale, g.ctypes.data))
        return g


def autotune(matcher: GpuMatcher, data_ptr: int, nbytes: int, repeats: int = 2,
             configs: int = 0) -> Tuple[np.ndarray, np.ndarray, int]:
    """Exhaustive sweep over the routing table on device-resident data
    (`configs`: bit mask of the entries to time, 0 = all)."""
    times = np.zeros(NCFG, dtype=np.float32)
    mask = np.zeros(NCFG, dtype=np.float32)
    best = ctypes.c_uint32(0)
    N.check(N.load().ws_autotune(matcher.ctx.handle, matcher.handle, ctypes.c_void_p(data_ptr), nbytes,
                                 N.MEM_DEVICE, N.MODE_OVERLAPPING, repeats, configs, times.ctypes.data,
                                 mask.ctypes.data, ctypes.byref(best)))
    return times, mask, int(best.value)


class AutoMatcher:
    """AutoMatcher (core/src/router.rs:59-291) on the GPU: every request is
    routed to a kernel configuration by the learned router when one is given,
    else by the default rule (the reference's single threshold has no kernel
    choice to make)."""

    def __init__(self, patterns: PatternSet, router: Optional[Router] = None,
                 config: Optional[AutoMatcherConfig] = None, context: Optional[Context] = None,
                 matches_per_mib: float = 0.0):
        self.matcher = GpuMatcher(patterns, config, context)
        self.router = router
        self.info = self.matcher.info()
        self.mask = supported_mask(self.info)
        self.matches_per_mib = matches_per_mib

    def route(self, nbytes: int) -> Tuning:
        if self.router is None:
            return Tuning()
        choice, _ = self.router.predict(features(self.info, nbytes, self.matches_per_mib), self.mask)
        return config_tuning(int(choice[0]))

    def scan(self, data, mode="overlapping"):
        buf = data if isinstance(data, (bytes, bytearray)) else bytes(data)
        return self.matcher.scan(buf, mode=mode, tuning=self.route(len(buf)))

    def scan_device(self, ptr: int, nbytes: int, out_ptr: int, cap: int) -> int:
        return self.matcher.scan_device(ptr, nbytes, out_ptr, cap, tuning=self.route(nbytes))


# --------------------------------------------------------------------------
# The request workload: 20 pattern sets and log-uniform request sizes.
# --------------------------------------------------------------------------

@dataclass
class PatternSetSpec:
    index: int
    kind: str              # "printable" or "kmer"
    n_patterns: int
    literal_fraction: float
    patterns: List[tuple] = field(default_factory=list)   # (kind, bytes)
    pieces: List[bytes] = field(default_factory=list)     # strings that match, for planting


def request_pattern_sets(seed: int = 5, count: int = 20, max_patterns: int = 10_000) -> List[PatternSetSpec]:
    """count sets with 10..max_patterns patterns (log-spaced) and literal
    fractions cycling 1, .75, .5, .25, 0; every other all-literal set is ACGT
    k-mers."""
    out = []
    rng = np.random.default_rng(seed)
    for i in range(count):
        n = int(round(10 * (max_patterns / 10) ** (i / max(1, count - 1))))
        frac = [1.0, 0.75, 0.5, 0.25, 0.0][i % 5]
        if frac == 1.0 and i % 2 == 1:
            lits = W.kmer_set(n, seed=seed * 100 + i)
            spec = PatternSetSpec(i, "kmer", n, 1.0, [(0, l) for l in lits], lits[: 512])
        else:
            nl = int(round(n * frac))
            lits = W.random_literals(nl, 4, 16, W.LIT_ALPHABET, seed=seed * 100 + i)
            specs = W.regex_set(n - nl, seed=seed * 100 + i + 50)
            pats = [(0, l) for l in lits] + [(1, s.source.encode()) for s in specs]
            pieces = lits[:256] + [s.sample(rng) for s in specs[:256]]
            spec = PatternSetSpec(i, "printable", n, frac, pats, pieces)
        out.append(spec)
    return out


def build_patternset(spec: PatternSetSpec) -> PatternSet:
    b = PatternSet.builder()
    for kind, p in spec.patterns:
        (b.literal if kind == 0 else b.regex)(p)
    return b.build()


def request_sizes(n: int, seed: int, lo: int = 256, hi: int = 256 << 20) -> np.ndarray:
    r = np.random.default_rng(seed)
    return np.exp(r.uniform(np.log(lo), np.log(hi), size=n)).astype(np.int64)


class DevicePools:
    """Device-resident input pools that requests slice: printable ASCII and
    ACGT, each at two planted-match densities (one piece per 4 KiB and one per
    256 KiB), so the match-density feature varies across requests."""

    GAPS = (4 << 10, 256 << 10)

    def __init__(self, specs: Sequence[PatternSetSpec], size: int = 256 << 20, seed: int = 7):
        from .device import DeviceBuffer

        printable_pieces = [p for s in specs if s.kind == "printable" for p in s.pieces[:64]]
        kmer_pieces = [p for s in specs if s.kind == "kmer" for p in s.pieces[:64]] or [b"ACGTACGTACGT"]
        self.dev, self.density = {}, {}
        for d, gap in enumerate(self.GAPS):
            pr = W.printable_input(size, seed + 10 * d)
            W.plant(pr, printable_pieces, gap, seed + 10 * d + 1)
            ac = W.kmer_input(size, kmer_pieces, seed=seed + 10 * d + 2, plant_gap=gap)
            for kind, arr in (("printable", pr), ("kmer", ac)):
                self.dev[(kind, d)] = DeviceBuffer.from_array(arr)
                self.density[(kind, d)] = (1 << 20) / gap
            del pr, ac
        self.size = size

    def ptr(self, kind: str, density: int, offset: int) -> int:
        return self.dev[(kind, density)].ptr + offset


@dataclass
class Request:
    rid: int
    set_index: int
    density: int
    offset: int
    nbytes: int


def make_requests(n: int, specs, pools: DevicePools, seed: int) -> List[Request]:
    r = np.random.default_rng(seed)
    sizes = np.minimum(request_sizes(n, seed), pools.size)
    sets = r.integers(0, len(specs), size=n)
    dens = r.integers(0, len(DevicePools.GAPS), size=n)
    offs = (r.random(n) * (pools.size - sizes + 1)).astype(np.int64)
    return [Request(i, int(sets[i]), int(dens[i]), int(offs[i]), int(sizes[i])) for i in range(n)]


def default_config_index(info: dict, nbytes: int) -> int:
    """The configuration the library picks without a router (WS_VARIANT_AUTO on
    a single scan, the device-sized pipeline): the LDS factor DFA at 4 blocks
    per CU when the hash window is at most 2 bytes, else the hash filter at 16
    positions per lane up to 16 MiB, 64 from 256 MiB and 32 between."""
    if info["hash_window"] <= 2 and info["variants_supported"] & (1 << N.VARIANT_DFA_LDS):
        return 13
    if nbytes <= (16 << 20):
        return 0
    return 2 if nbytes >= (256 << 20) else 1


def harvest(reqs, specs, pools, matchers, repeats: int = 2, subset: int = 0, seed: int = 0):
    """Autotune of each request: features, log-ms targets, masks, times and
    the default configuration. subset = 0 times the whole table (exhaustive);
    subset = k times k random entries per request (training data: the loss
    is masked, so each request teaches the configurations it timed)."""
    X, Y, M, T, D = [], [], [], [], []
    infos = {i: m.info() for i, m in matchers.items()} if isinstance(matchers, dict) else \
        {i: m.info() for i, m in enumerate(matchers)}
    rng = np.random.default_rng(seed)
    for q in reqs:
        spec = specs[q.set_index]
        m = matchers[q.set_index]
        configs = 0
        if subset:
            for c in rng.choice(NCFG, size=min(subset, NCFG), replace=False):
                configs |= 1 << int(c)
        times, mask, best = autotune(m, pools.ptr(spec.kind, q.density, q.offset), q.nbytes, repeats, configs)
        info = infos[q.set_index]
        X.append(features(info, q.nbytes, pools.density[(spec.kind, q.density)]))
        Y.append(np.where(mask > 0, np.log(np.maximum(times, 1e-3)), 0.0))
        M.append(mask)
        T.append(times)
        D.append(default_config_index(info, q.nbytes))
    return (np.array(X, np.float32), np.array(Y, np.float32), np.array(M, np.float32), np.array(T, np.float32),
            np.array(D, np.int64))


def evaluate(router: Router, X, T, M, D) -> Dict[str, float]:
    """Regret of the router's choice against the exhaustive best, per request,
    and of the default rule (no router), plus the total times."""
    choice, _ = router.predict(X, M)
    best = np.array([t[m > 0].min() for t, m in zip(T, M)])
    got = T[np.arange(len(T)), choice.astype(int)]
    regret = got / best - 1.0
    default = T[np.arange(len(T)), D]
    return {
        "median_regret": float(np.median(regret)),
        "mean_regret": float(np.mean(regret)),
        "p90_regret": float(np.quantile(regret, 0.9)),
        "default_median_regret": float(np.median(default / best - 1.0)),
        "router_total_ms": float(got.sum()),
        "default_total_ms": float(default.sum()),
        "best_total_ms": float(best.sum()),
        "exhaustive_tuning_ms": float(T.sum()),
    }


# --------------------------------------------------------------------------
# The whole study: harvest -> train -> regret -> serve.
# --------------------------------------------------------------------------

def _route_fn(router: Optional[Router], specs, pools, infos, masks):
    from .serve import Job  # noqa: F401  (type only)

    tunings = [config_tuning(i) for i in range(NCFG)]

    def route(jobs):
        if router is None:
            out = []
            for j in jobs:
                c = default_config_index(infos[j.patterns], j.nbytes)
                out.append((c, tunings[c]))
            return out
        X = np.stack([features(infos[j.patterns], j.nbytes, j.density) for j in jobs])
        M = np.stack([masks[j.patterns] for j in jobs])
        choice, _ = router.predict(X, M)
        return [(int(c), tunings[int(c)]) for c in choice]

    return route


def run_study(n_requests: int = 50_000, n_train: int = 1500, n_test: int = 600, pool_size: int = 256 << 20,
              slots: int = 4, load: float = 0.6, seed: int = 5, epochs: int = 300, n_sets: int = 20,
              max_patterns: int = 10_000, train_subset: int = NCFG // 2, log=None) -> dict:
    """Harvest autotuner timings, train the router on the GPU, measure its
    regret on held-out requests against the exhaustive sweep, and serve a
    Poisson stream of `n_requests` through the continuous-batching loop with
    the router on the request path."""
    from .serve import Job, RequestLoop, poisson_arrivals, summarize

    def say(msg):
        if log:
            log(msg)

    t = time.perf_counter()
    specs = request_pattern_sets(seed, n_sets, max_patterns)
    sets = [build_patternset(s) for s in specs]
    ctx = default_context(0)
    matchers = [GpuMatcher(ps, context=ctx) for ps in sets]
    infos = [m.info() for m in matchers]
    masks = [supported_mask(i) for i in infos]
    say(f"compiled {len(sets)} pattern sets in {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    pools = DevicePools(specs, pool_size, seed + 2)
    say(f"input pools ready in {time.perf_counter() - t:.1f}s")

    train = make_requests(n_train, specs, pools, seed + 11)
    test = make_requests(n_test, specs, pools, seed + 12)
    t = time.perf_counter()
    Xtr, Ytr, Mtr, Ttr, Dtr = harvest(train, specs, pools, matchers, subset=train_subset, seed=seed + 15)
    harvest_s = time.perf_counter() - t
    say(f"harvested {n_train} training requests x {train_subset or NCFG} of {NCFG} configs in {harvest_s:.1f}s")
    router = Router(seed=seed)
    t = time.perf_counter()
    loss = router.train(Xtr, Ytr, Mtr, epochs=epochs, batch=256, lr=2e-3, seed=seed)
    train_s = time.perf_counter() - t
    say(f"trained router in {train_s:.1f}s, loss {loss:.4f}")
    Xte, Yte, Mte, Tte, Dte = harvest(test, specs, pools, matchers)
    ev = evaluate(router, Xte, Tte, Mte, Dte)
    ev_train = evaluate(router, Xtr, Ttr, Mtr, Dtr)
    # Inference cost on the request path: one request, median over calls.
    lat = []
    for k in range(200):
        c0 = time.perf_counter()
        router.predict(Xte[k % len(Xte)], Mte[k % len(Mte)])
        lat.append(time.perf_counter() - c0)
    infer_ms = float(np.median(lat)) * 1e3
    sweep_ms = float((Tte * Mte).sum(axis=1).mean())   # one timed run of every supported config
    say(f"regret median {ev['median_regret']:.3f} (default {ev['default_median_regret']:.3f}), "
        f"infer {infer_ms:.3f} ms vs sweep {sweep_ms:.3f} ms")

    # Serving. First a burst (every request waiting at t=0) measures the
    # loop's capacity; then a Poisson stream offered at `load` x capacity
    # measures latency under continuous batching.
    reqs = make_requests(n_requests, specs, pools, seed + 13)

    def jobs_for(arrivals):
        out = []
        for q, a in zip(reqs, arrivals):
            kind = specs[q.set_index].kind
            out.append(Job(q.rid, q.set_index, pools.ptr(kind, q.density, q.offset), q.nbytes, float(a),
                           density=pools.density[(kind, q.density)]))
        return out

    loop = RequestLoop(sets, _route_fn(router, specs, pools, infos, masks), slots=slots)
    t = time.perf_counter()
    for s in loop.slots:
        for i in range(len(sets)):
            s.matcher(i).scan_device_array(pools.ptr(specs[i].kind, 0, 0), pool_size,
                                           tuning=config_tuning(default_config_index(infos[i], pool_size)))
    say(f"slots warmed in {time.perf_counter() - t:.1f}s")
    n_burst = min(n_requests, 5000)
    burst_jobs = jobs_for(np.zeros(n_requests))[:n_burst]
    burst = summarize(loop.run(burst_jobs), [j.nbytes for j in burst_jobs])
    say(f"burst of {n_burst}: {burst}")
    rate = load * burst["requests_per_s"]
    jobs = jobs_for(poisson_arrivals(n_requests, rate, seed + 14))
    say(f"serving {n_requests} requests at {rate:.0f}/s")
    outs = loop.run(jobs)
    loop.close()
    serving = summarize(outs, [j.nbytes for j in jobs])
    busy = sum(o.done - o.admitted for o in outs)
    serving["mean_concurrency"] = round(busy / max(1e-9, serving["wall_s"]), 2)
    serving.update({"slots": slots, "offered_rate_per_s": round(rate, 1), "load": load,
                    "matches": int(sum(o.count for o in outs)),
                    "capacity_requests_per_s": burst["requests_per_s"],
                    "capacity_gbs": burst["throughput_gbs"], "burst_requests": n_burst})
    say(f"served: {serving}")
    amortized = (n_requests * sweep_ms) / (harvest_s * 1e3 + train_s * 1e3 + n_requests * infer_ms)
    return {
        "router": {
            "median_regret": round(ev["median_regret"], 4),
            "mean_regret": round(ev["mean_regret"], 4),
            "p90_regret": round(ev["p90_regret"], 4),
            "train_median_regret": round(ev_train["median_regret"], 4),
            "no_router_median_regret": round(ev["default_median_regret"], 4),
            "router_total_ms": round(ev["router_total_ms"], 2),
            "no_router_total_ms": round(ev["default_total_ms"], 2),
            "best_total_ms": round(ev["best_total_ms"], 2),
            "infer_ms": round(infer_ms, 4),
            "sweep_ms_per_request": round(sweep_ms, 4),
            "tuning_cost_ratio": round(sweep_ms / infer_ms, 1),
            "tuning_cost_ratio_amortized": round(amortized, 1),
            "train_requests": n_train, "test_requests": n_test, "train_loss": round(loss, 5),
            "harvest_s": round(harvest_s, 2), "train_s": round(train_s, 2),
            "configs": NCFG, "train_configs_per_request": train_subset or NCFG, "pattern_sets": len(sets),
        },
        "serving": serving,
    }
