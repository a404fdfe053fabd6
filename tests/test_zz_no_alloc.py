"""The allocation-free entry point (ws_scan_noalloc) against ws_scan.

Through the C ABI: device-resident input and output, literal-only and
regex sets, every kernel variant; two calls in a row, the second on a
smaller input, each equal to ws_scan's device-sized result on the same
bytes (and that to the oracle). Inputs beyond its maximum are refused.
"""
import ctypes

import numpy as np
import pytest

from oracle.oracle import Oracle
from warpstate_amd import GpuMatcher, PatternSet, Tuning, VARIANT_DFA_GLOBAL, VARIANT_DFA_LDS, VARIANT_HASH
from warpstate_amd import _native as N
from warpstate_amd import workloads as W
from warpstate_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 16 << 20


def build(pats):
    b = PatternSet.builder()
    for kind, p in pats:
        (b.literal if kind == 0 else b.regex)(p)
    return b.build()


def scan_noalloc(m, dev, length, cap, tuning=None):
    out = DeviceBuffer(12 * max(cap, 1))
    count = ctypes.c_uint64(0)
    t = tuning.to_c() if tuning else None
    st = N.load().ws_scan_noalloc(m.ctx.handle, m.handle, ctypes.c_void_p(dev.ptr), length,
                                  ctypes.byref(t) if t else None, ctypes.c_void_p(out.ptr), cap,
                                  ctypes.byref(count))
    return st, int(count.value), out


def check_pair(pats, data, small, tuning=None):
    m = GpuMatcher(build(pats))
    dev = DeviceBuffer.from_array(data)
    for length in (len(data), small):
        print(f"noalloc len={length} pats={len(pats)} tuning={tuning}", flush=True)
        exp = m.scan_device_array(dev.ptr, length, tuning=tuning)
        assert np.array_equal(exp, Oracle(pats).scan(data[:length], 0))
        st, n, out = scan_noalloc(m, dev, length, len(exp) + 64, tuning)
        assert st == 0, N.load().ws_last_error()
        assert n == len(exp)
        got = out.download(np.uint8, 12 * n).view(exp.dtype)
        assert np.array_equal(got, exp)
    return len(exp)


def test_literal_set_two_calls():
    pats, data = W.literal_workload(size=16 << 20, seed=0)
    check_pair([(0, p) for p in pats], data, (5 << 20) + 3)


def test_regex_and_mixed_sets_every_variant():
    pats, _, specs = W.mixed_set(120, 40, seed=91)
    rng = np.random.default_rng(92)
    data = W.printable_input(3 << 20, 93)
    W.plant(data, [p for k, p in pats if k == 0] + [s.sample(rng) for s in specs], 400, 94)
    for v in (VARIANT_HASH, VARIANT_DFA_GLOBAL, VARIANT_DFA_LDS):
        check_pair(pats, data, 1_000_001, Tuning(variant=v))
    srcs, d2, _ = W.regex_workload(size=4 << 20, n=200, seed=2)
    check_pair([(1, s) for s in srcs], d2, 777_777)


def test_limits_are_errors():
    pats, data = W.literal_workload(size=NOALLOC_MAX + 16, seed=0)
    m = GpuMatcher(build([(0, p) for p in pats]))
    dev = DeviceBuffer.from_array(data)
    st, n, _ = scan_noalloc(m, dev, NOALLOC_MAX + 1, 1024)
    assert st == 10 and n == 0  # InputTooLarge
    # more matches than the caller's output holds: the required count
    st, n, _ = scan_noalloc(m, dev, NOALLOC_MAX, 4)
    exp = m.scan_device_array(dev.ptr, NOALLOC_MAX)
    assert st == 11 and n == len(exp) > 4
    # more literal matches than the fixed buffers hold: an error, no result
    one = np.frombuffer(b"ab" * (100_000), dtype=np.uint8)
    m2 = GpuMatcher(build([(0, b"ab"), (0, b"ba")]))
    dev2 = DeviceBuffer.from_array(one)
    st, n, _ = scan_noalloc(m2, dev2, len(one), 400_000)
    assert st == 11 and n == 0
    # and the next call works again
    assert check_pair([(0, p) for p in pats], data[: 1 << 20], 100_000) > 0
