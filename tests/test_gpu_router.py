"""Learned router on the GPU: forward pass and gradients of the HIP MLP kernels
against the CPU reference, training, and routing invariance of results."""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from warpstate_amd import GpuMatcher, PatternSet, Tuning
from warpstate_amd import workloads as W
from warpstate_amd.router import NCFG, Router, autotune, config_tuning, supported_mask

pytestmark = pytest.mark.gpu


def test_forward_and_gradient_match_cpu_reference():
    r = Router(seed=3)
    p = r.params().astype(np.float64)
    rng = np.random.default_rng(1)
    x = rng.normal(0, 1, (300, 8)).astype(np.float32)
    x[:, 7] = 0
    y = rng.normal(0, 1, (300, NCFG)).astype(np.float32)
    m = (rng.random((300, NCFG)) > 0.25).astype(np.float32)
    out = r.forward(x)
    ref, _ = M.forward(p, x)
    np.testing.assert_allclose(out, ref, rtol=1e-4, atol=1e-4)
    g = r.gradient(x, y, m, 1.0 / m.sum())
    gref = M.gradient(p, x, y, m, 1.0 / m.sum())
    # float32 accumulation over 300 samples: relative to the gradient scale
    np.testing.assert_allclose(g, gref, rtol=2e-3, atol=2e-5 * np.abs(gref).max())


def test_training_learns_a_routing_rule():
    # Synthetic costs: config 0 wins below feature0 = 0, config 3 above.
    rng = np.random.default_rng(2)
    x = rng.normal(0, 1, (2000, 8)).astype(np.float32)
    x[:, 7] = 0
    t = np.ones((2000, NCFG), np.float32) * 5.0
    t[:, 0] = 2.0 + np.maximum(x[:, 0], 0) * 4
    t[:, 3] = 2.0 + np.maximum(-x[:, 0], 0) * 4
    mask = np.ones((2000, NCFG), np.float32)
    r = Router(seed=4)
    loss = r.train(x, np.log(t), mask, epochs=60, batch=256, lr=3e-3)
    assert loss < 0.05
    choice, pred = r.predict(x, mask)
    want = np.where(x[:, 0] < 0, 0, 3)
    assert (choice == want).mean() > 0.95


def test_every_configuration_returns_the_same_matches():
    from warpstate_amd.device import DeviceBuffer

    pats, _, specs = W.mixed_set(120, 40, seed=12)
    b = PatternSet.builder()
    for kind, p in pats:
        (b.literal if kind == 0 else b.regex)(p)
    ps = b.build()
    data = W.printable_input(3 << 20, 5)
    W.plant(data, [p for _, p in pats[:120]] + [s.sample(np.random.default_rng(3)) for s in specs], 700, 6)
    m = GpuMatcher(ps)
    mask = supported_mask(m.info())
    ref = m.scan_array(data.tobytes(), tuning=Tuning())
    for i in range(NCFG):
        if mask[i]:
            got = m.scan_array(data.tobytes(), tuning=config_tuning(i))
            assert np.array_equal(got, ref), i
    dev = DeviceBuffer.from_array(data)
    times, tmask, best = autotune(m, dev.ptr, len(data))
    assert np.array_equal(tmask, mask)
    assert tmask[best] and times[best] == times[tmask > 0].min() and (times[tmask > 0] > 0).all()
    # A subset of the table: only the selected supported entries are timed.
    sel = 0b1010000000101001
    times2, tmask2, best2 = autotune(m, dev.ptr, len(data), configs=sel)
    want = mask * np.array([(sel >> i) & 1 for i in range(NCFG)], np.float32)
    assert np.array_equal(tmask2, want) and (times2[tmask2 == 0] == 0).all() and tmask2[best2]


def test_router_beats_the_default_rule_on_held_out_requests():
    """The method without the network is the library's fixed choice
    (default_config_index: the reference's AutoMatcher has one size threshold,
    core/src/router.rs:205-225). Trained on harvested timings of one set of
    requests, the router is scored on timings of other requests over the same
    20 request pattern sets (up to 2,000 patterns, sizes log-uniform 256 B -
    32 MiB). Thresholds: total time at most 97% of the default rule's, mean
    per-request speedup at least 1.05, median regret below the default's."""
    from warpstate_amd.router import (DevicePools, build_patternset, evaluate, harvest, make_requests,
                                      request_pattern_sets)
    from warpstate_amd.api import default_context

    specs = request_pattern_sets(seed=5, count=20, max_patterns=2000)
    ctx = default_context(0)
    matchers = [GpuMatcher(build_patternset(s), context=ctx) for s in specs]
    pools = DevicePools(specs, size=32 << 20, seed=9)
    train = make_requests(500, specs, pools, seed=21)
    test = make_requests(250, specs, pools, seed=22)
    Xtr, Ytr, Mtr, Ttr, Dtr = harvest(train, specs, pools, matchers)
    Xte, Yte, Mte, Tte, Dte = harvest(test, specs, pools, matchers)
    r = Router(seed=5)
    r.train(Xtr, Ytr, Mtr, epochs=300, batch=256, lr=2e-3, seed=5)
    ev = evaluate(r, Xte, Tte, Mte, Dte)
    choice, _ = r.predict(Xte, Mte)
    got = Tte[np.arange(len(Tte)), choice.astype(int)]
    default = Tte[np.arange(len(Tte)), Dte]
    speedup = float(np.mean(default / got))
    print(f"router {ev['router_total_ms']:.2f} ms, default {ev['default_total_ms']:.2f} ms, "
          f"best {ev['best_total_ms']:.2f} ms, mean speedup {speedup:.3f}, median regret "
          f"{ev['median_regret']:.4f} vs {ev['default_median_regret']:.4f}")
    assert ev["router_total_ms"] <= 0.97 * ev["default_total_ms"]
    assert speedup >= 1.05
    assert ev["median_regret"] < ev["default_median_regret"]
