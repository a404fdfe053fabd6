"""Router pieces that run without a GPU: the CPU MLP reference (pinned by finite
differences), the routing table and the request features."""
import ctypes

import numpy as np

from oracle import mlp_oracle as M
from warpstate_amd import PatternSet
from warpstate_amd import _native as N


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    flat = rng.normal(0, 0.3, M.N_PARAMS)
    x = rng.normal(0, 1, (6, M.IN))
    y = rng.normal(0, 1, (6, M.OUT))
    m = (rng.random((6, M.OUT)) > 0.3).astype(np.float64)
    g = M.gradient(flat, x, y, m, 0.1)
    idx = rng.choice(M.N_PARAMS, 60, replace=False)
    for i in idx:
        e = np.zeros(M.N_PARAMS)
        e[i] = 1e-6
        fd = (M.loss(flat + e, x, y, m, 0.1) - M.loss(flat - e, x, y, m, 0.1)) / 2e-6
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), i


def test_param_count_matches_layout():
    # 8 -> 64 -> 64 -> 64 -> 16
    assert N.load().ws_router_param_count() == M.N_PARAMS == 64 * 9 + 2 * 64 * 65 + 16 * 65 == 9936


def test_config_table():
    seen = set()
    for i in range(16):
        t = N.WsTuning()
        N.check(N.load().ws_config_tuning(i, ctypes.byref(t)))
        assert t.variant in (1, 2, 3) and t.flags in (0, N.TUNING_CHECKED)
        seen.add((t.variant, t.bytes_per_lane, t.blocks_per_cu, t.flags))
    assert len(seen) == 16
    t = N.WsTuning()
    assert N.load().ws_config_tuning(16, ctypes.byref(t)) == 13


def test_features():
    ps = PatternSet.builder().literal("needle").literal("haystack").regex("ab[0-9]{1,3}").build()
    info = ps.analyze()
    ci = N.WsPatternsInfo(**info)
    out = (ctypes.c_float * 8)()
    N.check(N.load().ws_router_features(ctypes.byref(ci), 1 << 20, 3.0, out))
    f = list(out)
    assert abs(f[0] - np.log2((1 << 20) + 1)) < 1e-4
    assert abs(f[1] - np.log2(3)) < 1e-5
    assert abs(f[4] - 2 / 3) < 1e-6 and f[5] == 6 and abs(f[6] - 2.0) < 1e-6 and f[7] == 0
