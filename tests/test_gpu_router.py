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
    y = rng.normal(0, 1, (300, 8)).astype(np.float32)
    m = (rng.random((300, 8)) > 0.25).astype(np.float32)
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
    t = np.ones((2000, 8), np.float32) * 5.0
    t[:, 0] = 2.0 + np.maximum(x[:, 0], 0) * 4
    t[:, 3] = 2.0 + np.maximum(-x[:, 0], 0) * 4
    mask = np.ones((2000, 8), np.float32)
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
