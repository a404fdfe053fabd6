"""HotSwapPatternSet (core/src/pattern/mod.rs:17-70) on the device: scans use
the active set, swap reports whether it changed anything (Arc::ptr_eq), and
scans racing with swaps each see exactly one set, never a mix."""
import threading

import numpy as np
import pytest

from oracle.oracle import Oracle
from warpstate_amd import Context, HotSwapPatternSet, PatternSet
from warpstate_amd import workloads as W

pytestmark = pytest.mark.gpu


def build(pats):
    b = PatternSet.builder()
    for kind, p in pats:
        (b.literal_bytes if kind == 0 else b.regex)(p)
    return b.build()


def sets():
    lits, data = W.literal_workload(size=1 << 18, seed=2)
    a = [(0, p) for p in lits[:32]]
    b_pats, _, _ = W.mixed_set(40, 10, seed=8)
    data = bytearray(data.tobytes())
    for k, (_, p) in enumerate(b_pats[:40]):  # plant set B's literals too
        o = 1000 + 5000 * k
        data[o:o + len(p)] = p
    return a, b_pats, bytes(data)


def test_hotswap_scans_active_set_and_swap_reports_change():
    a, b, data = sets()
    psa, psb = build(a), build(b)
    hs = HotSwapPatternSet(psa)
    exp_a, exp_b = Oracle(a).scan_list(data, 1), Oracle(b).scan_list(data, 1)
    assert exp_a != exp_b and exp_a and exp_b
    assert hs.scan(data) == exp_a
    assert hs.swap(psb) is True
    assert hs.scan(data) == exp_b
    assert hs.swap(psb) is False  # already active (Arc::ptr_eq)
    assert hs.swap(psa) is True
    assert hs.scan(data, mode="overlapping") == Oracle(a).scan_list(data, 0)
    hs.close()


def test_hotswap_racing_scans_see_one_set_each():
    a, b, data = sets()
    psa, psb = build(a), build(b)
    hs = HotSwapPatternSet(psa)
    exp = {tuple(Oracle(a).scan_list(data, 1)): "a", tuple(Oracle(b).scan_list(data, 1)): "b"}
    seen, errors = [], []
    stop = threading.Event()

    def scanner():
        ctx = Context(0)
        try:
            while not stop.is_set():
                got = tuple(hs.scan(data, ctx=ctx))
                if got not in exp:
                    errors.append(len(got))
                    return
                seen.append(exp[got])
        finally:
            ctx.close()

    t = threading.Thread(target=scanner)
    t.start()
    import time
    deadline = time.time() + 60
    for k in range(40):
        hs.swap(psb if k % 2 == 0 else psa)
        # let a scan or two run against each swap
        while len(seen) <= k and not errors and t.is_alive() and time.time() < deadline:
            time.sleep(0.001)
    stop.set()
    t.join(timeout=120)
    assert not t.is_alive()
    assert not errors, f"a scan saw neither set ({errors})"
    assert seen
    hs.close()
