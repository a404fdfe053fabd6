#!/usr/bin/env python3
"""Scan throughput benchmark on one MI355X.

Default workload (the headline): the k-mer literal scan — 10,000 ACGT
literals (lengths 12-32), 1 GiB of ACGT input with 0.5% 'N' bytes and one
planted k-mer per 4 KiB. A step is one whole ws_scan call over the
device-resident input: prefilter + verify kernels, ordering (sort) and
packing of the sorted match list into a device buffer. `value` is input bytes
per step time (GB/s).

Other workloads (--workload): `literal` (64 literals, 16 MiB printable; a
rotating pool of 32 input copies so every step reads HBM, not the 256 MiB
MALL), `regex` (200 grammar regexes, 1 GiB, StreamPipeline with 64 MiB
chunks and overlap = longest match), `batch` (100,000 log-normal items,
~1 GiB, one scan_batch call over device-resident items through a batch plan
made once, matches written to device memory).

`roofline` covers the dominant kernel (the rolling-hash prefilter,
k_hash_filter), timed with HIP events on the library's stream inside the timed
region. `cpu_baseline` is the C oracle on the box's host cores (byte-range
split over threads for the literal workloads). With the default workload the
learned-router study also runs (harvest, train, regret, continuous-batching
request loop) and its results ride in the same JSON line.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "scan GB/s on 1 MI355X, % of HBM peak; learned-router regret vs autotune"


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def profile_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE, calibrated against a copy kernel)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.txt")
    try:
        with open(path) as f:
            m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read(), re.M)
        return float(m.group(1)) if m else None
    except OSError:
        return None


def cpu_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def build(patterns):
    from warpstate_amd import PatternSet

    b = PatternSet.builder()
    for kind, p in patterns:
        (b.literal if kind == 0 else b.regex)(p)
    return b.build()


def cpu_rate(patterns, data: np.ndarray, threads: int, mode: int = 0):
    """One oracle pass over `data`; returns (GB/s, seconds)."""
    from oracle.oracle import MATCH_DTYPE, Oracle

    o = Oracle(patterns)
    out = np.zeros(max(1 << 16, len(data) // 256), dtype=MATCH_DTYPE)
    c0 = time.perf_counter()
    n = o.scan_into(data, out, mode, threads)
    dt = time.perf_counter() - c0
    if n > len(out):  # the count pass is complete; the timing stands
        pass
    return len(data) / dt / 1e9, dt


class Workload:
    def __init__(self, args):
        from warpstate_amd import AutoMatcherConfig, Context, GpuMatcher, StreamPipeline, Tuning
        from warpstate_amd import workloads as W
        from warpstate_amd.device import DeviceBuffer

        self.args = args
        self.name = args.workload
        variant = {"auto": 0, "hash": 1, "dfa_global": 2, "dfa_lds": 3}[args.variant]
        self.tuning = Tuning(variant=variant, bytes_per_lane=args.bytes_per_lane, blocks_per_cu=args.blocks_per_cu)
        t0 = time.time()
        self.copies = 1
        if self.name == "kmer":
            lits, data = W.kmer_workload(size=args.size or (1 << 30), n=10_000, seed=1)
            self.patterns = [(0, p) for p in lits]
            self.config = {"workload": "kmer-literal-scan", "patterns": len(lits), "pattern_len": "12-32 ACGT",
                           "input_bytes": len(data), "n_frac": 0.005, "plant_every": 4096, "mode": "overlapping"}
        elif self.name == "literal":
            lits, data = W.literal_workload(size=args.size or (16 << 20), seed=0)
            self.patterns = [(0, p) for p in lits]
            self.copies = 32
            self.config = {"workload": "literal-scan", "patterns": len(lits), "pattern_len": "4-16 [a-z_]",
                           "input_bytes": len(data), "plant_every": 65536, "mode": "overlapping",
                           "input_copies": self.copies}
        elif self.name == "regex":
            srcs, data, maxlen = W.regex_workload(size=args.size or (1 << 30), n=200, seed=2)
            self.patterns = [(1, s.encode()) for s in srcs]
            self.maxlen = maxlen
            self.config = {"workload": "regex-stream-scan", "patterns": len(srcs), "input_bytes": len(data),
                           "chunk": 64 << 20, "overlap": maxlen, "plant_every": 32768, "mode": "overlapping"}
        elif self.name == "batch":
            pats, lits, specs = W.mixed_set(700, 300, seed=3)
            pieces = lits[:256] + [s.sample(np.random.default_rng(5)) for s in specs[:256]]
            ids, offs, sizes, data = W.batch_items(100_000, seed=4, pieces=pieces)
            self.patterns = pats
            self.ids, self.offs, self.sizes = ids, offs, sizes
            self.config = {"workload": "uneven-batch-scan", "patterns": len(pats), "items": len(ids),
                           "input_bytes": int(sizes.sum()), "item_median": 4096, "sigma": 1.5,
                           "mode": "overlapping"}
        else:
            raise SystemExit(f"unknown workload {self.name}")
        self.data = data
        log(f"generated {len(self.patterns)} patterns, {len(data)} bytes in {time.time() - t0:.1f}s")
        self.ps = build(self.patterns)
        self.ctx = Context(0, AutoMatcherConfig())
        self.m = GpuMatcher(self.ps, context=self.ctx)
        self.dev = [DeviceBuffer.from_array(data)]
        for _ in range(self.copies - 1):
            b = DeviceBuffer(len(data))
            b.copy_from(self.dev[0])
            self.dev.append(b)
        self.cap = 16 << 20
        self.out = DeviceBuffer(self.cap * 12)
        self.n = len(data) if self.name != "batch" else int(self.sizes.sum())
        self.k = 0
        if self.name in ("kmer", "literal"):
            # One prepared foreign call per input copy (arguments converted once).
            self.calls = [self.m.prepare_device(b.ptr, self.n, self.out.ptr, self.cap, tuning=self.tuning)
                          for b in self.dev]
        if self.name == "regex":
            self.pipe = StreamPipeline(self.m, self.maxlen)
        if self.name == "batch":
            from warpstate_amd import BatchPlan

            self.ptrs = (self.dev[0].ptr + self.offs).astype(np.uint64)
            # The batch's host planning (id order, segments, work units,
            # dense positions) is done once, outside the timed steps; its
            # cost is reported as plan_ms, and an unplanned call as
            # unplanned_ms.
            t1 = time.perf_counter()
            self.plan = BatchPlan(self.ctx, self.ids, self.ptrs, self.sizes)
            self.plan_ms = (time.perf_counter() - t1) * 1e3
            self.tagged_cap = 4 << 20
            self.tagged_out = DeviceBuffer(self.tagged_cap * 24)

    def step(self) -> int:
        k = self.k % self.copies
        buf = self.dev[k]
        self.k += 1
        if self.name in ("kmer", "literal"):
            return self.calls[k]()
        if self.name == "regex":
            return self.pipe.scan_device(buf.ptr, self.n, self.out.ptr, self.cap, tuning=self.tuning,
                                         chunk_size=64 << 20)
        from warpstate_amd import scan_batch_planned

        return scan_batch_planned(self.m, self.plan, tuning=self.tuning, out_ptr=self.tagged_out.ptr,
                                  capacity=self.tagged_cap)

    def unplanned_ms(self, reps: int = 5) -> float:
        """One scan_batch call over the same device items with its host
        planning and a host result array (best of `reps`)."""
        from warpstate_amd import scan_batch_device

        best = float("inf")
        for _ in range(reps):
            t = time.perf_counter()
            scan_batch_device(self.m, self.ids, self.ptrs, self.sizes, tuning=self.tuning)
            best = min(best, time.perf_counter() - t)
        return best * 1e3

    def cpu_baseline(self):
        threads = cpu_threads()
        if self.name in ("kmer", "literal"):
            sample = self.data if self.name == "literal" else self.data[: self.args.cpu_sample or len(self.data)]
            rate_mt, dt_mt = cpu_rate(self.patterns, sample, threads)
            one = sample[: min(len(sample), 64 << 20)]
            rate_1, dt_1 = cpu_rate(self.patterns, one, 1)
            return {"value": round(rate_mt, 4), "unit": "GB/s", "cores": threads, "kind": "port",
                    "sample": f"{len(sample) >> 20} MiB of the same input, overlapping literal scan, C oracle "
                              f"split over {threads} threads by byte range ({dt_mt:.1f} s)",
                    "single_core_value": round(rate_1, 4),
                    "single_core_sample": f"first {len(one) >> 20} MiB, one thread ({dt_1:.1f} s)"}
        # Regex / mixed sets: the oracle's regex engine is single-threaded.
        size = self.args.cpu_sample or (8 << 20)
        sample = self.data[:size]
        rate_1, dt_1 = cpu_rate(self.patterns, sample, 1)
        return {"value": round(rate_1, 4), "unit": "GB/s", "cores": 1, "kind": "port",
                "sample": f"first {len(sample) >> 20} MiB of the same input as one scan, overlapping literals + "
                          f"leftmost-first regexes, C oracle, one thread ({dt_1:.1f} s)"}


def index_timing(w) -> dict:
    """Loading the workload's set from a version-4 compiled index against
    compiling it from its patterns; both include the device upload."""
    import ctypes

    from warpstate_amd import CompiledPatternIndex
    from warpstate_amd import _native as N

    lib = N.load()
    data = CompiledPatternIndex.build(w.ps)
    arr, _keep = w.ps._c_patterns()
    best = {"compile": float("inf"), "load": float("inf")}
    for _ in range(3):
        for what in ("compile", "load"):
            h = ctypes.c_void_p()
            t = time.perf_counter()
            if what == "compile":
                st = lib.ws_patterns_compile(w.ctx.handle, arr, len(w.ps.patterns), 0, ctypes.byref(h))
            else:
                st = lib.ws_index_load(w.ctx.handle, data, len(data), ctypes.byref(h))
            dt = time.perf_counter() - t
            N.check(st)
            lib.ws_patterns_free(h)
            best[what] = min(best[what], dt * 1e3)
    return {"index_bytes": len(data), "compile_ms": round(best["compile"], 2),
            "index_load_ms": round(best["load"], 2), "load_speedup": round(best["compile"] / best["load"], 1)}


def copy_calibration(args) -> int:
    """1 GiB device-to-device copies, for calibrating FETCH_SIZE / WRITE_SIZE."""
    from warpstate_amd.device import DeviceBuffer, synchronize

    n = args.size or (1 << 30)
    a, b = DeviceBuffer(n), DeviceBuffer(n)
    a.zero()
    for _ in range(args.warmup):
        b.copy_from(a)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.copy_from(a)
    synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    print(json.dumps({"copy_bytes": n, "ms_per_copy": round(dt * 1e3, 4),
                      "gbs_read_plus_write": round(2 * n / dt / 1e9, 1)}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # Defaults long enough for the GPU's clocks to settle: with 3 warm-up
    # and 20 timed k-mer steps (8 ms) the first steps ran 10-15% slower.
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=100)
    # Untimed steps for a fixed wall time before the warm-up: the GPU's
    # clocks ramp for about a second under load, so a short run (say 5
    # warm-up and 20 timed steps, 8 ms of k-mer work) would otherwise time
    # steps that run 10-15% slower than settled ones.
    ap.add_argument("--settle-s", type=float, default=3.0)
    ap.add_argument("--workload", default="kmer", choices=["kmer", "literal", "regex", "batch"])
    ap.add_argument("--size", type=int, default=0, help="input bytes (default: the workload's own)")
    ap.add_argument("--variant", default="auto", choices=["auto", "hash", "dfa_global", "dfa_lds"])
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--bytes-per-lane", type=int, default=0, help="hash filter positions per lane (16/32/64)")
    ap.add_argument("--cpu-sample", type=int, default=0, help="bytes for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--router", dest="router", action="store_true", default=None,
                    help="run the learned-router study (default: on for the kmer workload)")
    ap.add_argument("--no-router", dest="router", action="store_false")
    ap.add_argument("--requests", type=int, default=50_000, help="requests in the serving loop")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every call eagerly (no hipGraph capture), e.g. under a profiler")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; multi-GPU runs are not part of this build", file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration(args)

    from warpstate_amd.device import synchronize

    w = Workload(args)
    if args.no_graphs:
        w.ctx.set_graphs(False)
    settle_steps = 0
    t_settle = time.perf_counter()
    while time.perf_counter() - t_settle < args.settle_s:
        w.step()
        settle_steps += 1
    settle_s = time.perf_counter() - t_settle
    log(f"clock settle: {settle_steps} untimed steps in {settle_s:.2f}s")
    for _ in range(args.warmup):
        count = w.step()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        count = w.step()
    synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    # Per-step kernel timings (the filter's device-clock stamps) from the
    # same number of steps run again with profiling on, so that neither the
    # profiling's host work (events around each call) nor reading the
    # timings back adds to the timed region. (Profiled calls are captured
    # as graphs of their own: a few untimed steps first.)
    w.ctx.set_profiling(True)
    for _ in range(min(args.warmup, 5) + 2):
        w.step()
    kern_ms, ver_ms, order_ms, rx_ms, hits = [], [], [], [], 0
    for _ in range(args.steps):
        w.step()
        t = w.ctx.last_timings()
        kern_ms.append(t.prefilter_ms)
        ver_ms.append(t.verify_ms)
        order_ms.append(t.order_ms)
        rx_ms.append(t.regex_ms)
        hits = t.hits
    value = w.n / (ms_per_step * 1e-3) / 1e9
    k_ms = float(np.mean(kern_ms))
    # Algorithmic bytes of a step: the input read once, each match written
    # once (12-byte Match; 24-byte TaggedMatch in a batch), 16 bytes per
    # batch item descriptor.
    n_items = int(w.config.get("items", 0))
    alg_bytes = w.n + (24 if n_items else 12) * int(count) + 16 * n_items
    # The stream workload launches the filter once per chunk; kernel_ms and
    # the algorithmic bytes cover the whole step, traffic one launch.
    launches = -(-w.n // w.config["chunk"]) if "chunk" in w.config else 1
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    info = w.m.info()
    # The prefilter kernel's time comes from its own device-clock stamps,
    # valid in every step; the other phases' events are re-recorded only by
    # eager launches (-1 once the step replays a captured graph).
    breakdown = {"prefilter": round(k_ms, 4)}
    for name, vals in (("verify", ver_ms), ("regex", rx_ms), ("order", order_ms)):
        if min(vals) >= 0:
            breakdown[name] = round(float(np.mean(vals)), 4)
    if len(breakdown) == 1:
        breakdown["other"] = round(ms_per_step - k_ms, 4)
    log(f"matches={count} hits={hits} kernel_ms={k_ms:.4f} breakdown={breakdown} step_ms={ms_per_step:.4f} "
        f"variants={info['variants_supported']} K={info['hash_window']}")

    cpu = None if args.no_cpu else w.cpu_baseline()
    if cpu:
        log(f"cpu baseline {cpu}")
    traffic = profile_traffic("k_hash_filter", w.name)
    config = dict(w.config)
    config.update({"variant": args.variant, "matches": int(count), "parallelism": "single-gpu"})
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": config,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_hash_filter", "kernel_ms": round(k_ms, 4),
                     "algorithmic_bytes": alg_bytes, "launches_per_step": launches,
                     "traffic_per": "launch"},
        "cpu_baseline": cpu,
        "pct_hbm_peak_whole_step": round(100 * value / HBM_PEAK_GBS, 2),
        "graphs_replayed": w.ctx.graph_count,
        "settle_s": round(settle_s, 2),
        "settle_steps": settle_steps,
        "step_breakdown_ms": breakdown,
    }
    if w.name == "batch":
        line["batch"] = {"plan_ms": round(w.plan_ms, 3), "unplanned_call_ms": round(w.unplanned_ms(), 3),
                         "output": "device"}
        log(f"batch {line['batch']}")
    line["compiled_index"] = index_timing(w)
    log(f"compiled index {line['compiled_index']}")
    run_router = args.router if args.router is not None else (w.name == "kmer")
    if run_router:
        from warpstate_amd.router import run_study

        del w
        t1 = time.time()
        study = run_study(n_requests=args.requests, log=lambda m: log(f"{time.time() - t1:.1f}s {m}"))
        line.update(study)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
