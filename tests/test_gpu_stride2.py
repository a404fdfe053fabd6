"""GPU parity of the stride-2 rolling-hash filter against the oracle.

Sets whose shortest factor has K + 1 or more bytes, for an odd window K
(5, 7, 11, 15), hash the windows at every second position only; each factor
sits in the filter and the verification table with its two grams (offsets
0 and 1), and a hit at window position p checks the starts p and p - 1.
These cases pin what that changes: factors of exactly K + 1 bytes at both
parities and at the edges of the input, segment starts of either parity
(batch items, unaligned device input), the last start a stream chunk emits
(a start p - 1 whose window p is the chunk's emit bound), case folding,
regex factors, and buckets crowded by hundreds of equal grams.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from warpstate_amd import GpuMatcher, ScanItem, StreamPipeline, Tuning, VARIANT_HASH, scan_batch, scan_batch_device
from warpstate_amd import workloads as W
from warpstate_amd.device import DeviceBuffer

from test_gpu_parity import as_list, assert_same, build, check_all_variants

pytestmark = pytest.mark.gpu


def window_of(pats, ci=False):
    return build(pats, ci).analyze()["hash_window"]


def planted(size, pats, seed, alpha=None, gap=700):
    r = np.random.default_rng(seed)
    alpha = W.LIT_ALPHABET if alpha is None else alpha
    data = alpha[r.integers(0, len(alpha), size=size)]
    W.plant(data, [p for _, p in pats], gap, seed + 1)
    return data


def test_long_literals_every_lane_width():
    lits = W.random_literals(300, 6, 20, W.LIT_ALPHABET, seed=21)
    pats = [(0, p) for p in lits]
    assert window_of(pats) == 5
    data = planted(1 << 20, pats, 22).tobytes()
    exp = check_all_variants(pats, data)
    assert len(exp) > 1000


def test_factors_of_window_plus_one_at_both_parities_and_edges():
    # 12-byte literals (window 11): every start parity, the input's first
    # and last possible starts.
    r = np.random.default_rng(31)
    lits = [W.ACGT[r.integers(0, 4, size=12)].tobytes() for _ in range(40)]
    lits += [W.ACGT[r.integers(0, 4, size=int(n))].tobytes() for n in r.integers(13, 30, size=40)]
    pats = [(0, p) for p in lits]
    assert window_of(pats) == 11
    for size in (4096, 4097, 70_001):
        data = bytearray(W.ACGT[r.integers(0, 4, size=size)].tobytes())
        for k, start in enumerate([0, 1, 2, 3, size - 12, size - 13, size - 14] + list(range(17, size - 40, 97))):
            lit = lits[k % 40]
            data[start:start + 12] = lit
        data = bytes(data)
        print(f"window+1 size={size}", flush=True)
        check_all_variants(pats, data)
        # the same bytes at both parities of the segment start
        m = GpuMatcher(build(pats))
        dev = DeviceBuffer.from_array(np.frombuffer(b"x" + data, dtype=np.uint8))
        for off in (0, 1):
            src = (b"x" + data)[off:]
            got = m.scan_device_array(dev.ptr + off, len(src))
            assert_same(as_list(got), Oracle(pats).scan_list(src), f"device offset {off}")


def test_case_insensitive_long_literals():
    lits = [b"NeedleInHay", b"HAYSTACKS", b"QuickBrownFox", b"lazy_dogs"]
    pats = [(0, p) for p in lits]
    assert window_of(pats, ci=True) == 7
    r = np.random.default_rng(41)
    words = [b"needleinhay", b"NEEDLEINHAY", b"haystacks", b"HaYsTaCkS", b"quickbrownfox", b"LAZY_DOGS", b"lazy_dog"]
    data = b" ".join(words[int(i)] for i in r.integers(0, len(words), size=4000))
    check_all_variants(pats, data, mode=0, ci=True)


def test_regex_factors_and_literals():
    # Regex factors of 6+ bytes: candidate starts p and p - 1 go to the walk.
    pats = [(1, "abcdef[0-9]{1,4}"), (1, "qwerty[a-z]+x"), (1, "(zxcvbn|poiuyt)_[a-f]{2}"), (0, b"lkjhgfd")]
    assert window_of(pats) == 5
    r = np.random.default_rng(51)
    pieces = [b"abcdef12", b"abcdef9999x", b"qwertyabcx", b"qwertyx", b"zxcvbn_af", b"poiuyt_zz", b"lkjhgfd",
              b"abcde", b"qwert"]
    data = bytearray(W.LIT_ALPHABET[r.integers(0, len(W.LIT_ALPHABET), size=300_000)].tobytes())
    for pos in r.integers(0, len(data) - 16, size=3000):
        p = pieces[int(r.integers(0, len(pieces)))]
        data[int(pos):int(pos) + len(p)] = p
    exp = check_all_variants(pats, bytes(data))
    assert len(exp) > 1000


def test_crowded_buckets_of_equal_grams():
    # 300 literals share their first 13 bytes: both grams of every one hash
    # alike, so one bucket holds 300 entries per offset (the overflow list).
    r = np.random.default_rng(61)
    head = b"ACGTTGCAACGTA"
    lits = sorted({head + W.ACGT[r.integers(0, 4, size=int(n))].tobytes() for n in r.integers(1, 8, size=300)})
    pats = [(0, p) for p in lits]
    assert window_of(pats) == 11
    data = planted(200_000, pats, 62, alpha=W.ACGT, gap=300).tobytes()
    exp = check_all_variants(pats, data)
    assert len(exp) > 300


@pytest.mark.parametrize("filler,window", [(0, 11), (1200, 12)])
def test_buckets_of_three_to_nine_equal_grams(filler, window):
    # Groups of 3..9 literals share their first 12 bytes, so a bucket holds
    # each group's equal keys: a crowded bucket's slot lists the keys of its
    # first five further entries (several of them equal here) and reads the
    # rest one by one. 1,200 more random 12-32 byte literals make the
    # 12-byte window selective enough for stride 1.
    r = np.random.default_rng(81)
    lits = set()
    for g in range(3, 10):
        head = W.ACGT[r.integers(0, 4, size=12)].tobytes()
        tails = set()
        while len(tails) < g:
            tails.add(W.ACGT[r.integers(0, 4, size=int(r.integers(1, 6)))].tobytes())
        lits |= {head + t for t in tails}
    lits |= set(W.random_literals(filler, 12, 32, W.ACGT, seed=82)) if filler else set()
    pats = [(0, p) for p in sorted(lits)]
    assert window_of(pats) == window
    data = planted(300_000, pats, 83, alpha=W.ACGT, gap=200).tobytes()
    exp = check_all_variants(pats, data)
    assert len(exp) > 1000


@pytest.mark.parametrize("chunk", [4096, 4097, 65_537])
def test_stream_chunk_edges(chunk):
    # A factor starting at the last emitted position of a chunk has its
    # second window on the chunk's emit bound.
    r = np.random.default_rng(71)
    lits = [W.LIT_ALPHABET[r.integers(0, len(W.LIT_ALPHABET), size=int(n))].tobytes()
            for n in r.integers(8, 14, size=50)]
    pats = [(0, p) for p in lits]
    assert window_of(pats) == 7
    data = bytearray(planted(600_000, pats, 72).tobytes())
    for k, edge in enumerate(range(chunk, len(data) - 20, chunk)):
        lit = lits[k % len(lits)]
        start = edge - 1 - (k % 3)  # the chunk's last start, one before, two before
        data[start:start + len(lit)] = lit
    data = bytes(data)
    m = GpuMatcher(build(pats))
    exp = Oracle(pats).scan_list(data)
    sp = StreamPipeline(m, 16)
    for t in (None, Tuning(checked=True), Tuning(variant=VARIANT_HASH, bytes_per_lane=16)):
        assert_same(as_list(sp.scan_array(data, chunk_size=chunk, tuning=t)), exp, f"stream {chunk} {t}")


def test_batch_items_of_both_parities():
    r = np.random.default_rng(81)
    lits = W.random_literals(200, 6, 18, W.LIT_ALPHABET, seed=82)
    pats = [(0, p) for p in lits]
    ids, offs, sizes, buf = W.batch_items(600, seed=83, pieces=lits, plant_gap=200)
    items = [ScanItem(int(i), buf[o:o + s].tobytes()) for i, o, s in zip(ids, offs, sizes)]
    m = GpuMatcher(build(pats))
    o = Oracle(pats)
    exp = [(it.id, x) for it in items for x in o.scan_list(it.data)]
    assert len(exp) > 500
    assert_same([(t.source_id, tuple(t.matched)) for t in scan_batch(m, items)], exp, "batch")
    dev = DeviceBuffer.from_array(buf)
    got = scan_batch_device(m, ids.astype(np.uint64), (dev.ptr + offs).astype(np.uint64), sizes)
    got_l = [(int(a), (int(b), int(c), int(d))) for a, b, c, d in
             zip(got["source_id"], got["pattern_id"], got["start"], got["end"])]
    exp_sorted = sorted(exp, key=lambda t: (t[0], t[1][1], t[1][0], t[1][2]))
    assert_same(got_l, exp_sorted, "device batch")
    assert sorted(set(int(x) % 2 for x in offs)) == [0, 1]
