"""ws_scan_async / ws_poll / ws_wait: GpuMatcher::scan's async contract
(core/src/gpu/mod.rs:223-260) through the C ABI. Every collected result is
exactly ws_scan's (and the oracle's) for the same arguments."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from warpstate_amd import (MATCH_DTYPE, AutoMatcherConfig, Context, GpuMatcher, PatternSet, Tuning,
                           WarpstateError)
from warpstate_amd import workloads as W
from warpstate_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu


def build(pats):
    b = PatternSet.builder()
    for kind, p in pats:
        (b.literal if kind == 0 else b.regex)(p)
    return b.build()


def mixed(seed, size):
    pats, _, specs = W.mixed_set(120, 40, seed=seed)
    rng = np.random.default_rng(seed + 1)
    data = W.printable_input(size, seed + 2)
    W.plant(data, [p for k, p in pats if k == 0] + [s.sample(rng) for s in specs], 400, seed + 3)
    return pats, data


def collect(out, n):
    return out.download(MATCH_DTYPE, n) if n else np.zeros(0, MATCH_DTYPE)


def test_async_equals_sync_and_oracle_on_repeats():
    pats, data = mixed(61, 3 << 20)
    ctx = Context(0, AutoMatcherConfig())
    m = GpuMatcher(build(pats), context=ctx)
    exp = Oracle(pats).scan(data, 0)
    dev = DeviceBuffer.from_array(data)
    out = DeviceBuffer(len(exp) * MATCH_DTYPE.itemsize + 4096)
    cap = len(exp) + 100
    # Repeats: eager, then captured, then replayed graphs, all asynchronous.
    for k in range(5):
        pend = m.scan_device_async(dev.ptr, len(data), out.ptr, cap)
        n = pend.wait()
        assert pend.ready()
        assert np.array_equal(collect(out, n), exp), k
    assert ctx.graph_count >= 1
    # Host output, and polling until ready.
    host = np.zeros(cap, MATCH_DTYPE)
    pend = m.scan_device_async(dev.ptr, len(data), host.ctypes.data, cap, out_on_device=False)
    while not pend.ready():
        pass
    n = pend.wait()
    assert np.array_equal(host[:n], exp)


def test_other_calls_complete_the_outstanding_one():
    pats, data = mixed(71, 2 << 20)
    _, other = mixed(72, 1 << 20)
    m = GpuMatcher(build(pats))
    o = Oracle(pats)
    exp, exp_other = o.scan(data, 0), o.scan(other, 0)
    dev = DeviceBuffer.from_array(data)
    out = DeviceBuffer((len(exp) + 64) * MATCH_DTYPE.itemsize)
    pend = m.scan_device_async(dev.ptr, len(data), out.ptr, len(exp) + 64)
    # A second async call before wait is refused; the first stays outstanding.
    with pytest.raises(WarpstateError):
        m.scan_device_async(dev.ptr, len(data), out.ptr, len(exp) + 64)
    # A synchronous scan on the same context runs after it.
    assert np.array_equal(m.scan_array(other), exp_other)
    n = pend.wait()
    assert np.array_equal(collect(out, n), exp)
    # Nothing outstanding: ws_wait is an error.
    from warpstate_amd import _native as N
    import ctypes
    c = ctypes.c_uint64(0)
    assert N.load().ws_wait(m.ctx.handle, ctypes.byref(c)) != 0


def test_async_errors_and_redo_arrive_at_wait():
    # Too small an output: the overflow status and required count come from wait.
    pats, data = mixed(81, 1 << 20)
    m = GpuMatcher(build(pats))
    exp = Oracle(pats).scan(data, 0)
    assert len(exp) > 20
    dev = DeviceBuffer.from_array(data)
    out = DeviceBuffer(4096)
    pend = m.scan_device_async(dev.ptr, len(data), out.ptr, 10)
    with pytest.raises(WarpstateError) as e:
        pend.wait()
    assert e.value.count == len(exp)
    # Device-sized buffers overflow (500,000 regex candidates): redone on the
    # checked pipeline inside wait, same answer as ws_scan.
    data2 = np.frombuffer(b"ab" * 500_000 + b"ab7", np.uint8)
    m2 = GpuMatcher(build([(1, b"ab[0-9]")]))
    dev2 = DeviceBuffer.from_array(data2)
    out2 = DeviceBuffer(4096)
    n = m2.scan_device_async(dev2.ptr, len(data2), out2.ptr, 64).wait()
    got = collect(out2, n)
    assert [(int(a), int(b), int(c)) for a, b, c in zip(got["pattern_id"], got["start"], got["end"])] == \
        [(0, 1_000_000, 1_000_003)]


def test_async_standard_mode_and_checked_tuning_finish_at_once():
    alpha = np.frombuffer(b"ab", dtype=np.uint8)
    lits = W.random_literals(64, 2, 8, alpha, seed=91)
    data = alpha[np.random.default_rng(91).integers(0, 2, size=200_000)]
    pats = [(0, l) for l in lits]
    m = GpuMatcher(build(pats))
    o = Oracle(pats)
    dev = DeviceBuffer.from_array(data)
    for mode, tuning in (("standard", None), ("overlapping", Tuning(checked=True))):
        exp = o.scan(data.tobytes(), 1 if mode == "standard" else 0)
        out = DeviceBuffer((len(exp) + 16) * MATCH_DTYPE.itemsize)
        pend = m.scan_device_async(dev.ptr, len(data), out.ptr, len(exp) + 16, mode=mode, tuning=tuning)
        assert pend.ready()
        n = pend.wait()
        assert np.array_equal(collect(out, n), exp), mode
