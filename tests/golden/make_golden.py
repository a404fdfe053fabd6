#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/.

kats.json — known-answer cases whose expected outputs are stated in the
  reference's own tests (file:line cited per case; paths relative to the
  reference checkout). Inputs and outputs are data transcribed from those
  assertions.
regex_vectors.json — leftmost-first find_iter vectors computed with Python's
  `re` (bytes). For the supported syntax (literal bytes, classes, `.`, bounded
  and unbounded greedy/lazy repeats, alternation, groups, no anchors) `re`
  and the reference's regex crate with multi_line(true).unicode(false) both
  implement backtracking-priority (leftmost-first) semantics with `.`
  excluding only '\\n'.
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

L, R = 0, 1

# (name, source, patterns [(kind, text)], case_insensitive, input, mode, expected [(pid, start, end)])
KATS = [
    ("overlap_abc_bc", "core/tests/correctness_guarantees.rs:15-41",
     [(L, "abc"), (L, "bc")], False, "abc", "overlapping", [(0, 0, 3), (1, 1, 3)]),
    ("dupes_overlapping", "core/tests/correctness_guarantees.rs:282-300",
     [(L, "DUPE"), (L, "DUPE")], False, "xxDUPEyy", "overlapping", [(0, 2, 6), (1, 2, 6)]),
    ("dupes_standard", "core/tests/correctness_guarantees.rs:303-322",
     [(L, "DUPE"), (L, "DUPE")], False, "xxDUPEyy", "standard", [(0, 2, 6)]),
    ("all_null_overlapping", "core/tests/correctness_guarantees.rs:257-279",
     [(L, "\x00\x00\x00")], False, "\x00" * 5, "overlapping", [(0, 0, 3), (0, 1, 4), (0, 2, 5)]),
    ("null_in_pattern", "core/tests/correctness_guarantees.rs:229-242",
     [(L, "a\x00b")], False, "x a\x00b y", "standard", [(0, 2, 5)]),
    ("null_in_input", "core/tests/correctness_guarantees.rs:245-254",
     [(L, "end")], False, "abc\x00end\x00def", "standard", [(0, 4, 7)]),
    ("invalid_utf8_literal", "core/tests/correctness_guarantees.rs:195-210",
     [(L, "test")], False, "test\xfftest", "standard", [(0, 0, 4), (0, 5, 9)]),
    ("regex_raw_byte", "core/tests/correctness_guarantees.rs:213-222",
     [(R, r"\xFF")], False, "\x00\xff\x00", "standard", [(0, 1, 2)]),
    ("aa_standard", "core/src/cpu/mod.rs:352-364",
     [(L, "aa")], False, "aaaa", "standard", [(0, 0, 2), (0, 2, 4)]),
    ("aa_overlapping", "core/src/cpu/mod.rs:367-379",
     [(L, "aa")], False, "aaaa", "overlapping", [(0, 0, 2), (0, 1, 3), (0, 2, 4)]),
    ("single_memchr", "core/src/specialize.rs:670-681",
     [(L, "needle")], False, "needle in a haystack with needle", "standard", [(0, 0, 6), (0, 26, 32)]),
    ("multi_memchr", "core/src/specialize.rs:683-702",
     [(L, "foo"), (L, "bar"), (L, "baz")], False, "foo bar baz foo", "standard",
     [(0, 0, 3), (1, 4, 7), (2, 8, 11), (0, 12, 15)]),
    ("hash_scanner_small", "core/src/hash_scan.rs:531-542",
     [(L, "needle"), (L, "token"), (L, "secret")], False, "xxneedlexxsecretxxtoken", "standard",
     [(0, 2, 8), (2, 10, 16), (1, 18, 23)]),
    ("ci_hello", "core/tests/unit/harness_regex.rs:28-31",
     [(R, "(?i:hello)")], False, "Hello hello HELLO hElLo", "standard",
     [(0, 0, 5), (0, 6, 11), (0, 12, 17), (0, 18, 23)]),
    ("ci_needle_lines", "core/tests/unit/harness_regex.rs:34-40",
     [(R, "(?i:needle)")], False, "Needle\nneedle\nNEEDLE\n", "standard",
     [(0, 0, 6), (0, 7, 13), (0, 14, 20)]),
    ("ci_single_char", "core/tests/unit/harness_regex.rs:43-46",
     [(R, "(?i:a)")], False, "aAbBaA", "standard", [(0, 0, 1), (0, 1, 2), (0, 4, 5), (0, 5, 6)]),
    ("ci_no_match", "core/tests/unit/harness_regex.rs:49-52",
     [(R, "(?i:xyz)")], False, "hello world", "standard", []),
    ("digit_class", "core/tests/unit/harness_regex.rs:65-68",
     [(R, "[0-9]+")], False, "abc 123 def 456", "standard", [(0, 4, 7), (0, 12, 15)]),
    ("alternation", "core/tests/unit/harness_regex.rs:77-80",
     [(R, "foo|bar")], False, "foo baz bar", "standard", [(0, 0, 3), (0, 8, 11)]),
    ("colou_r", "core/tests/unit/harness_regex.rs:105-108",
     [(R, "colou?r")], False, "color colour", "standard", [(0, 0, 5), (0, 6, 12)]),
    ("non_capturing_group", "core/tests/unit/harness_regex.rs:171-174",
     [(R, "(?:foo|bar)baz")], False, "foobaz barbaz", "standard", [(0, 0, 6), (0, 7, 13)]),
    ("ci_regression_lines", "core/tests/unit/harness_regex.rs:266-274",
     [(R, "(?i:hello)")], False, "Hello\nhello\nHELLO", "standard", [(0, 0, 5), (0, 6, 11), (0, 12, 17)]),
    ("consecutive_ab", "core/tests/unit/harness_regex.rs:277-281",
     [(R, "ab")], False, "ababab", "standard", [(0, 0, 2), (0, 2, 4), (0, 4, 6)]),
    ("match_at_end", "core/tests/unit/harness_regex.rs:284-288",
     [(R, "end")], False, "the end", "standard", [(0, 4, 7)]),
    ("match_at_start", "core/tests/unit/harness_regex.rs:291-295",
     [(R, "start")], False, "start here", "standard", [(0, 0, 5)]),
    ("single_char_repeated", "core/tests/unit/harness_regex.rs:298-305",
     [(R, "x")], False, "xxxx", "standard", [(0, 0, 1), (0, 1, 2), (0, 2, 3), (0, 3, 4)]),
]

REGEX_PATTERNS = [
    "[0-9]+", "a[bc]*d", "x.{2,5}?y", "(foo|foobar)baz", "(?i:hello)", "[a-z]{3}[0-9]{1,3}",
    r"\d{2}-\d{2}", r"\w+@\w+", "(ab|a)(bc|c)", "q[^q\n]{0,4}q", "abc[a-z0-9]{1,8}[A-F0-9]{1,8}",
    "zz.{1,8}|yy[A-F0-9]{1,8}", "(?:ha){2,3}", "k[ey]+?s",
]


def bytes_of(s: str) -> bytes:
    return s.encode("latin-1")


def main() -> None:
    kats = []
    for name, src, pats, ci, inp, mode, exp in KATS:
        kats.append({"name": name, "source": src, "patterns": [[k, bytes_of(p).hex()] for k, p in pats],
                     "case_insensitive": ci, "input": bytes_of(inp).hex(), "mode": mode,
                     "expected": [list(e) for e in exp]})
    with open(os.path.join(HERE, "kats.json"), "w") as f:
        json.dump(kats, f, indent=1)

    from warpstate_amd import workloads as W

    vectors = []
    r = np.random.default_rng(1234)
    alpha = np.frombuffer(b"abcdxyzqk0123456789-@ABCDEFhlo\nfrbs_ey", dtype=np.uint8)
    for i, pat in enumerate(REGEX_PATTERNS):
        data = alpha[r.integers(0, len(alpha), size=600)].tobytes()
        found = [(m.start(), m.end()) for m in re.finditer(pat.encode(), data)]
        vectors.append({"pattern": pat, "input": data.hex(), "expected": found})
    specs = W.regex_set(12, seed=77)
    for s in specs:
        rng = np.random.default_rng(len(vectors))
        data = bytearray(W.printable_input(1500, len(vectors)).tobytes())
        for _ in range(12):
            piece = s.sample(rng)
            at = int(rng.integers(0, len(data) - len(piece)))
            data[at:at + len(piece)] = piece
        data = bytes(data)
        found = [(m.start(), m.end()) for m in re.finditer(s.source.encode(), data)]
        vectors.append({"pattern": s.source, "input": data.hex(), "expected": found})
    with open(os.path.join(HERE, "regex_vectors.json"), "w") as f:
        json.dump(vectors, f, indent=0)
    print(f"wrote {len(kats)} KATs and {len(vectors)} regex vectors")


if __name__ == "__main__":
    main()
