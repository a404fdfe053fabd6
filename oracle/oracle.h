/*
 * CPU oracle for the warpstate scan path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU semantics, used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker. The
 * product library never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference checkout). Match records are {pattern_id, start, end}, sorted
 * by (start, pattern_id, end) as core/src/cpu/mod.rs:67-78 does.
 */
#ifndef WS_ORACLE_H
#define WS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

typedef struct { uint32_t pattern_id, start, end; } or_match;

typedef struct {
  uint32_t kind;        /* 0 literal, 1 regex */
  const uint8_t* bytes;
  uint64_t len;
} or_pattern;

/* literal_prefilter_hash: FNV-1a over the first max(1, window) bytes
 * (core/src/pattern/compiler.rs:343-350). */
uint32_t or_fnv1a_window(const uint8_t* p, uint64_t len, uint32_t window);

/* LiteralPrefilterTable::build + probe (core/src/literal_prefilter.rs:23-119).
 * Writes the table into caller buffers sized by or_prefilter_sizes; returns 0. */
void or_prefilter_sizes(const uint32_t* lens, uint32_t n, uint32_t window,
                        uint32_t* n_buckets, uint32_t* n_entries);
int or_prefilter_build(const uint32_t* lens, const uint32_t* hashes, uint32_t n, uint32_t window,
                       uint32_t* prefix_meta /* window*4 */, uint32_t* bucket_ranges /* 2*nb */,
                       uint32_t* entries /* 2*n */);
/* Literal indices whose [hash, index] entry sits in the probed bucket with an
 * equal hash; returns how many were written (up to cap). */
uint32_t or_prefilter_probe(const uint32_t* prefix_meta, uint32_t window,
                            const uint32_t* bucket_ranges, uint32_t n_buckets,
                            const uint32_t* entries, uint32_t prefix_len, uint32_t hash,
                            uint32_t* out, uint32_t cap);

/* Compiled pattern set (literals + regex programs). */
typedef struct or_set or_set;
/* Returns NULL on error and writes a message to err (err_cap bytes). Status
 * codes follow the ws_status numbering (6 compile failure, 12 pathological,
 * 2 empty pattern, 1 empty set). */
or_set* or_compile(const or_pattern* pats, uint32_t n, int case_insensitive, int* status,
                   char* err, size_t err_cap);
void or_free(or_set* s);
uint32_t or_literal_count(const or_set* s);
uint32_t or_regex_count(const or_set* s);
/* Literal-reduction strategy that ScanStrategy::select picks: 1 leftmost-longest,
 * 2 multi-memchr, 3 hash-shortest, 0 none. */
uint32_t or_strategy(const or_set* s);

/* PatternSet::scan_overlapping (core/src/cpu/scan.rs:235-247,300-349). mode 0.
 * PatternSet::scan (core/src/pattern/mod.rs:112-140). mode 1.
 * Writes up to cap matches, returns the total count (may exceed cap). */
uint64_t or_scan(const or_set* s, const uint8_t* data, uint64_t len, int mode, or_match* out,
                 uint64_t cap);

/* scan_overlapping split over `threads` threads by byte range with a
 * (max_len - 1) halo (core/src/multi_gpu.rs:170-200); literal-only sets,
 * UINT64_MAX otherwise. Same result as or_scan(mode 0). */
uint64_t or_scan_overlapping_mt(const or_set* s, const uint8_t* data, uint64_t len, uint32_t threads,
                                or_match* out, uint64_t cap);

/* or_scan (either mode) over `threads` threads: literal byte-range shards
 * (mode 0) or one literal job (mode 1), and one job per regex. Same result. */
uint64_t or_scan_mt(const or_set* s, const uint8_t* data, uint64_t len, int mode, uint32_t threads,
                    or_match* out, uint64_t cap);

/* Leftmost-first find_iter of one regex (core/src/dfa/scan.rs:433-457). */
uint64_t or_regex_find_iter(const or_set* s, uint32_t regex_index, const uint8_t* data,
                            uint64_t len, or_match* out, uint64_t cap);

/* Anchored leftmost-first match of one regex at `start`; returns end or -1. */
int64_t or_regex_match_at(const or_set* s, uint32_t regex_index, const uint8_t* data,
                          uint64_t len, uint64_t start);

#endif
