"""The library's anchored regex DFAs, checked on the host (no GPU).

The DFA tables of a compiled set come out of its version-5 index (built
host-only by ws_index_build). A Python restatement of the device's use of
them -- the walk of k_rx_walk / k_rx_scan (start state by the byte before,
the match flag of a transition meaning "a match ends before this byte", the
end-of-text column) and the chain rule of k_chain_walk (greedy, an empty
match where the previous one ended skipped) -- must give the oracle's
find_iter on every regex vector, the reference's KATs and a seeded grammar
of assertions and empty-capable repeats. A wrong start context, a missed
delayed match or an end-of-text slip fails here before any GPU run.
"""
import ctypes
import json
import os
import re
import struct

import numpy as np
import pytest

from oracle.index_format import parse as parse_index
from oracle.oracle import Oracle
from warpstate_amd import _native as N

HERE = os.path.dirname(os.path.abspath(__file__))
DFA_MATCH, DFA_DEAD, DFA_MASK = 0x80000000, 0x40000000, 0x3FFFFFFF


def index_of(pats):
    lib = N.load()
    arr = (N.WsPattern * len(pats))()
    keep = []
    for i, (k, p) in enumerate(pats):
        b = p.encode() if isinstance(p, str) else bytes(p)
        keep.append(ctypes.create_string_buffer(b, len(b)))
        arr[i].kind, arr[i].bytes, arr[i].len, arr[i].name = k, ctypes.cast(keep[-1], ctypes.c_void_p), len(b), None
    n = ctypes.c_uint64(0)
    N.check(lib.ws_index_build(arr, len(pats), 0, 0, None, 0, ctypes.byref(n)))
    out = ctypes.create_string_buffer(n.value)
    N.check(lib.ws_index_build(arr, len(pats), 0, 0, out, n.value, ctypes.byref(n)))
    return out.raw[: n.value]


def dfas_of(pats):
    """The regex DFA sections of the set's version-5 index, in regex order."""
    p = parse_index(index_of(pats))
    assert p["version"] == 5
    t, off, out = p["tail"], 0, []
    for _ in range(p["regex_count"]):
        nstates, nclass = struct.unpack_from("<II", t, off)
        start = struct.unpack_from("<4I", t, off + 8)
        cmap = t[off + 24: off + 280]
        n = struct.unpack_from("<I", t, off + 280)[0]
        table = np.frombuffer(t, dtype="<u4", count=n, offset=off + 284)
        off += 284 + 4 * n
        assert n == nstates * (nclass + 1)
        out.append((nstates, nclass, start, cmap, table))
    return out


def ctx_of(b):
    if b == 10:
        return 1
    return 2 if (48 <= b <= 57 or 65 <= b <= 90 or 97 <= b <= 122 or b == 95) else 3


def match_at(d, data, st):
    _, nclass, start, cmap, table = d
    w = nclass + 1
    s = start[ctx_of(data[st - 1]) if st > 0 else 0]
    last = None
    for p in range(st, len(data)):
        e = int(table[s * w + cmap[data[p]]])
        if e & DFA_MATCH:
            last = p
        if e & DFA_DEAD:
            return last
        s = e & DFA_MASK
    if int(table[s * w + nclass]) & DFA_MATCH:
        last = len(data)
    return last


def device_find_iter(d, data):
    """Every position a candidate, its anchored match, then the chain."""
    out, last_end, have = [], 0, False
    for st in range(len(data) + 1):
        e = match_at(d, data, st)
        if e is None or st < last_end or (st == e and have and st == last_end):
            continue
        out.append((st, e))
        last_end, have = e, True
    return out


def check(pattern, data):
    dfas = dfas_of([(1, pattern)])
    if not dfas:  # a plain-literal regex compiles to a literal
        return
    assert device_find_iter(dfas[0], data) == Oracle([(1, pattern)]).find_iter(0, data), (pattern, data)


def load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


@pytest.mark.parametrize("vec", load("regex_vectors.json"), ids=lambda v: v["pattern"][:24])
def test_dfa_walk_and_chain_equal_oracle_on_vectors(vec):
    data = bytes.fromhex(vec["input"])
    check(vec["pattern"], data)
    dfas = dfas_of([(1, vec["pattern"])])
    if dfas:
        assert device_find_iter(dfas[0], data) == [tuple(e) for e in vec["expected"]]


@pytest.mark.parametrize("kat", [k for k in load("kats.json") if all(p[0] == 1 for p in k["patterns"])],
                         ids=lambda k: k["name"])
def test_dfa_walk_on_reference_regex_kats(kat):
    data = bytes.fromhex(kat["input"])
    pats = [(1, bytes.fromhex(h).decode("latin-1")) for _, h in kat["patterns"]]
    if kat["case_insensitive"]:
        pytest.skip("set-level case folding is not a regex DFA property")
    got = []
    dfas = dfas_of(pats)
    rx = 0
    for pid, (_, p) in enumerate(pats):
        if not dfas_of([(1, p)]):
            continue  # compiled as a literal
        got += [(pid, s, e) for s, e in device_find_iter(dfas[rx], data)]
        rx += 1
    if rx == len(pats):
        got.sort(key=lambda t: (t[1], t[0], t[2]))
        assert got == [tuple(e) for e in kat["expected"]], kat["source"]


@pytest.mark.parametrize("seed", range(3))
def test_dfa_walk_on_look_grammar(seed):
    from test_oracle import _look_regex

    r = np.random.default_rng(700 + seed)
    alpha = np.frombuffer(b"abfox1 \n_-", dtype=np.uint8)
    for _ in range(40):
        pat = _look_regex(r)
        data = alpha[r.integers(0, len(alpha), size=int(r.integers(0, 120)))].tobytes()
        check(pat, data)


def test_grammar_regexes_keep_their_prefilter_and_dfa_size():
    # The benchmark grammar (literal prefix, bounded atoms, alternation) has
    # no assertions: its DFAs keep one context, so the four start states are
    # one state and the table gains only the end-of-text column.
    from warpstate_amd import workloads as W

    specs = W.regex_set(20, seed=2)
    for d, s in zip(dfas_of([(1, s.source) for s in specs]), specs):
        assert len(set(d[2])) == 1, s.source
