"""Literal-only ordering on one block (k_lit_order_small). A context turns it
on after a call with at most 1,024 matches per run, a run with more than
4,096 is redone on the checked path and turns it off, and a later sparse call
turns it back on. Every call, on either path, equals the oracle exactly;
the sequences repeat device pointers so that captured graphs replay."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from warpstate_amd import Context, GpuMatcher, PatternSet, StreamPipeline
from warpstate_amd import workloads as W
from warpstate_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu


def literal_set_and_inputs():
    pats = W.literal_set(64, seed=3)
    b = PatternSet.builder()
    for p in pats:
        b.literal(p)
    inputs = {}
    for name, gap, seed in (("sparse", 16 << 10, 11), ("medium", 2 << 10, 13), ("dense", 256, 15)):
        d = W.printable_input(4 << 20, seed)
        W.plant(d, pats, gap, seed + 1)
        inputs[name] = d
    return pats, b.build(), inputs


def test_switches_with_the_match_count():
    pats, ps, inputs = literal_set_and_inputs()
    orc = Oracle([(0, p) for p in pats])
    exp = {k: orc.scan(v) for k, v in inputs.items()}
    assert len(exp["sparse"]) < 1024 < len(exp["medium"]) < 4096 < len(exp["dense"])
    dev = {k: DeviceBuffer.from_array(v) for k, v in inputs.items()}
    m = GpuMatcher(ps, context=Context(0))
    order = ["sparse", "sparse", "sparse", "dense", "dense", "medium", "sparse", "sparse", "medium", "medium",
             "dense", "sparse", "sparse"]
    for step, name in enumerate(order):
        print(f"step {step}: {name}", flush=True)
        got = m.scan_device_array(dev[name].ptr, len(inputs[name]))
        assert np.array_equal(got, exp[name]), (step, name, len(got), len(exp[name]))


def test_stream_runs():
    pats, ps, inputs = literal_set_and_inputs()
    orc = Oracle([(0, p) for p in pats])
    m = GpuMatcher(ps, context=Context(0))
    sp = StreamPipeline(m, 16)
    for name in ("sparse", "sparse", "medium", "sparse", "sparse", "dense", "sparse"):
        data = inputs[name]
        got = sp.scan_array(data, chunk_size=512 << 10)
        assert np.array_equal(got, orc.scan(data)), name


def test_prepared_call_equals_scan_device():
    from warpstate_amd import MATCH_DTYPE, WarpstateError

    pats, ps, inputs = literal_set_and_inputs()
    m = GpuMatcher(ps, context=Context(0))
    data = inputs["medium"]
    dev = DeviceBuffer.from_array(data)
    cap = 1 << 16
    out = DeviceBuffer(cap * 12)
    exp = Oracle([(0, p) for p in pats]).scan(data)
    call = m.prepare_device(dev.ptr, len(data), out.ptr, cap)
    for _ in range(3):
        n = call()
        assert n == len(exp)
        got = out.download(MATCH_DTYPE, n)
        assert np.array_equal(got, exp)
    small = m.prepare_device(dev.ptr, len(data), out.ptr, 16)
    with pytest.raises(WarpstateError) as e:
        small()
    assert e.value.status == 11
