"""The C-ABI library: loads, exports every declared symbol, validates on the host.

No compute calls here (this suite runs without a GPU).
"""
import ctypes
import os
import re

import pytest

from warpstate_amd import PatternSet, WarpstateError
from warpstate_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "warpstate.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:[\w\*]+\s+)+\**(ws_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(N.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    # the Python binding declares a signature for each of them
    assert set(syms) == set(N.SIGNATURES)


def test_abi_version():
    assert N.load().ws_abi_version() == 1


def test_struct_layouts_match_reference_records():
    # Match is 12 bytes (core/src/dfa/mod.rs:416-419); TaggedMatch carries a u64 id.
    assert N.MATCH_DTYPE.itemsize == 12
    assert N.TAGGED_DTYPE.itemsize == 24
    assert ctypes.sizeof(N.WsPattern) == 32
    assert ctypes.sizeof(N.WsItem) == 24


def test_no_device_reports_no_gpu_adapter():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    st = N.load().ws_ctx_create(0, None, ctypes.byref(h))
    assert st == 3 and not h.value


@pytest.mark.parametrize("pat,variant", [
    ("(a+)+", "PathologicalRegex"), ("([0-9]+)*", "PathologicalRegex"), ("a(", "PatternCompilationFailed"),
    ("a{3", "PatternCompilationFailed"), ("^*", "PatternCompilationFailed"),
])
def test_compile_errors_map_to_reference_variants(pat, variant):
    with pytest.raises(WarpstateError) as e:
        PatternSet.builder().regex(pat).build()
    assert e.value.variant == variant


def test_empty_pattern_and_set():
    with pytest.raises(WarpstateError) as e:
        PatternSet.builder().literal("").build()
    assert e.value.variant == "EmptyPattern"
    with pytest.raises(WarpstateError) as e:
        PatternSet.builder().build()
    assert e.value.variant == "EmptyPatternSet"


def test_analyze_reports_set_facts():
    from warpstate_amd import workloads as W

    b = PatternSet.builder()
    for p in W.kmer_set(2000, seed=1):
        b.literal(p)
    info = b.build().analyze()
    assert info["literal_count"] == 2000 and info["regex_count"] == 0
    # DNA grams of 11 bytes at every second position would be too frequent
    # for 2000 literals over 4 letters: 12-byte windows at every position
    assert info["min_literal_len"] >= 12 and info["hash_window"] == 12
    b = PatternSet.builder()
    for p in W.kmer_set(200, seed=1):
        b.literal(p)
    assert b.build().analyze()["hash_window"] == 11
    assert info["variants_supported"] & (1 << N.VARIANT_HASH)
    ps = PatternSet.builder().regex("abc[a-z0-9]{1,8}|de[A-F0-9]{1,8}").literal("abcd").build()
    info = ps.analyze()
    assert info["regex_count"] == 1 and info["literal_count"] == 1
    assert info["max_pattern_len"] == 11
    # a pure-literal regex is compiled as a literal (core/src/pattern/compiler.rs:179-212)
    assert PatternSet.builder().regex("needle").build().analyze()["literal_count"] == 1
