"""Seeded synthetic workloads for the scan benchmarks (numpy PCG64 generators).

Each generator returns plain Python/numpy data, so the same bytes reach the
GPU path and the CPU oracle. Sizes are parameters: tests use small ones, the
benchmark the full ones.

  literal_set / printable_input   64 literals [a-z_]{4..16}, printable text,
                                  one planted copy per 64 KiB on average
  kmer_set / kmer_input           ACGT literals {12..32}, ACGT text with 0.5% N,
                                  one planted k-mer per 4 KiB on average
  regex_set / regex_input         prefix + 1..3 atoms {[a-z0-9],[A-F0-9],.}{1,8},
                                  2..4-way top-level alternation
  batch_items                     log-normal item sizes (median 4 KiB, sigma 1.5)
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

PRINTABLE_LO, PRINTABLE_HI = 0x20, 0x7F  # [0x20, 0x7E]
LIT_ALPHABET = np.frombuffer(b"abcdefghijklmnopqrstuvwxyz_", dtype=np.uint8)
ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)
ATOMS = ("[a-z0-9]", "[A-F0-9]", ".")
_ATOM_BYTES = {
    "[a-z0-9]": np.frombuffer(b"abcdefghijklmnopqrstuvwxyz0123456789", dtype=np.uint8),
    "[A-F0-9]": np.frombuffer(b"ABCDEF0123456789", dtype=np.uint8),
    ".": np.arange(PRINTABLE_LO, PRINTABLE_HI, dtype=np.uint8),
}


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def random_literals(n: int, lmin: int, lmax: int, alphabet: np.ndarray, seed: int) -> List[bytes]:
    r = _rng(seed)
    lens = r.integers(lmin, lmax + 1, size=n)
    return [alphabet[r.integers(0, len(alphabet), size=int(L))].tobytes() for L in lens]


def plant(data: np.ndarray, pieces: List[bytes], mean_gap: int, seed: int) -> int:
    """Copies a random piece at geometric gaps (mean `mean_gap`); returns count."""
    r = _rng(seed)
    pos, n = 0, 0
    size = len(data)
    while True:
        pos += int(r.geometric(1.0 / mean_gap))
        piece = pieces[int(r.integers(0, len(pieces)))]
        if pos + len(piece) > size:
            return n
        data[pos:pos + len(piece)] = np.frombuffer(piece, dtype=np.uint8)
        pos += len(piece)
        n += 1


def literal_set(n: int = 64, seed: int = 0) -> List[bytes]:
    """64 literals, lengths uniform 4-16 over [a-z_], seed 0."""
    return random_literals(n, 4, 16, LIT_ALPHABET, seed)


def printable_input(size: int, seed: int) -> np.ndarray:
    return _rng(seed).integers(PRINTABLE_LO, PRINTABLE_HI, size=size, dtype=np.uint8)


def literal_workload(size: int = 16 << 20, seed: int = 0) -> Tuple[List[bytes], np.ndarray]:
    pats = literal_set(64, seed)
    data = printable_input(size, seed + 100)
    plant(data, pats, 64 * 1024, seed + 200)
    return pats, data


def kmer_set(n: int = 10_000, seed: int = 1) -> List[bytes]:
    """k-mer literals over ACGT, lengths uniform 12-32."""
    return random_literals(n, 12, 32, ACGT, seed)


def kmer_input(size: int, pats: List[bytes], seed: int = 1, n_frac: float = 0.005,
               plant_gap: int = 4096) -> np.ndarray:
    r = _rng(seed + 100)
    data = ACGT[r.integers(0, 4, size=size, dtype=np.uint8)]
    n_count = int(size * n_frac)
    if n_count:
        data[r.integers(0, size, size=n_count)] = ord("N")
    plant(data, pats, plant_gap, seed + 200)
    return data


def kmer_workload(size: int = 1 << 30, n: int = 10_000, seed: int = 1) -> Tuple[List[bytes], np.ndarray]:
    pats = kmer_set(n, seed)
    return pats, kmer_input(size, pats, seed)


@dataclass
class RegexSpec:
    """A generated pattern: branches of (prefix, [atoms])."""

    branches: List[Tuple[bytes, List[str]]]

    @property
    def source(self) -> str:
        return "|".join(p.decode() + "".join(a + "{1,8}" for a in atoms) for p, atoms in self.branches)

    @property
    def max_len(self) -> int:
        return max(len(p) + 8 * len(a) for p, a in self.branches)

    def sample(self, r: np.random.Generator) -> bytes:
        prefix, atoms = self.branches[int(r.integers(0, len(self.branches)))]
        out = bytearray(prefix)
        for a in atoms:
            alpha = _ATOM_BYTES[a]
            out += alpha[r.integers(0, len(alpha), size=int(r.integers(1, 9)))].tobytes()
        return bytes(out)


def regex_set(n: int = 200, seed: int = 2) -> List[RegexSpec]:
    """Prefix of 3-8 bytes over [a-z_], then 1-3 atoms, 2-4-way alternation."""
    r = _rng(seed)
    specs = []
    for _ in range(n):
        branches = []
        for _b in range(int(r.integers(2, 5))):
            plen = int(r.integers(3, 9))
            prefix = LIT_ALPHABET[r.integers(0, len(LIT_ALPHABET), size=plen)].tobytes()
            atoms = [ATOMS[int(r.integers(0, 3))] for _a in range(int(r.integers(1, 4)))]
            branches.append((prefix, atoms))
        specs.append(RegexSpec(branches))
    return specs


def regex_input(size: int, specs: List[RegexSpec], seed: int = 2, plant_gap: int = 32 * 1024) -> np.ndarray:
    data = printable_input(size, seed + 100)
    r = _rng(seed + 200)
    pieces = [specs[int(r.integers(0, len(specs)))].sample(r) for _ in range(4096)]
    plant(data, pieces, plant_gap, seed + 300)
    return data


def regex_workload(size: int = 1 << 30, n: int = 200, seed: int = 2):
    specs = regex_set(n, seed)
    return [s.source for s in specs], regex_input(size, specs, seed), max(s.max_len for s in specs)


def mixed_set(n_lit: int = 700, n_rx: int = 300, seed: int = 3):
    """700 literals 4-24 bytes + 300 regexes of the grammar above: (kind, bytes)."""
    lits = random_literals(n_lit, 4, 24, LIT_ALPHABET, seed)
    specs = regex_set(n_rx, seed + 1)
    pats = [(0, l) for l in lits] + [(1, s.source.encode()) for s in specs]
    return pats, lits, specs


def batch_sizes(n: int, seed: int, median: int = 4096, sigma: float = 1.5, lo: int = 64,
                hi: int = 4 << 20) -> np.ndarray:
    r = _rng(seed)
    s = np.exp(r.normal(np.log(median), sigma, size=n))
    return np.clip(np.round(s), lo, hi).astype(np.int64)


def batch_items(n: int = 100_000, seed: int = 4, pieces: List[bytes] | None = None,
                plant_gap: int = 32 * 1024):
    """(ids, offsets, lengths, buffer) with items laid out back to back."""
    sizes = batch_sizes(n, seed)
    total = int(sizes.sum())
    buf = printable_input(total, seed + 100)
    if pieces:
        plant(buf, pieces, plant_gap, seed + 200)
    offs = np.zeros(n, dtype=np.int64)
    offs[1:] = np.cumsum(sizes)[:-1]
    ids = np.arange(n, dtype=np.uint64)
    return ids, offs, sizes, buf
