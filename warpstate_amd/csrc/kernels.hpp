// Device-side parameter blocks shared by the kernels and the host runtime.
#pragma once
#include <cstdint>

#include <hip/hip_runtime.h>

namespace ws {

constexpr int kTPB = 256;                 // threads per block for the scan kernels
constexpr uint32_t kHashUnit = kTPB * 16; // bytes per hash-kernel work unit at 16 positions per lane
constexpr uint32_t kRecCap = 256;         // per-block LDS staging records (each list)
constexpr int kFilterTPB = 1024;          // threads per block of the streaming hash filter
// Largest span one scan covers (a single input, a stream, or a batch's items
// in one allocation): 4 GiB - 1 MiB, so that the filter's 32-bit bounds --
// a half unit's end (tile base + unit + 16) and a segment's rounded end --
// cannot wrap.
constexpr uint64_t kMaxScanBytes = (1ull << 32) - (1ull << 20);

// Match / candidate record. Match: {pattern_id, start, end, item}.
// Regex candidate: {regex index, start, 0, item}. Offsets are relative to the
// item (segment) start.
using Rec = uint4;

struct ScanCommon {
  const uint8_t* base;  // 16-byte aligned base of the scanned buffer
  uint64_t load_end;    // bytes [0, load_end) may be loaded; multiple of 16
  const uint4* units;   // {tile_base, item, segment start, segment end} per work unit, or null (one segment)
  uint32_t nunits;
  const uint2* items;   // {lo, hi} segment bounds in base coordinates
  uint32_t emit_end;    // only starts (relative to lo) below this are emitted
  uint32_t ci;          // ASCII case folding
  // Factors: verification strings and owners.
  const uint8_t* fbytes;
  const uint32_t* foff;
  const uint32_t* flen;
  const uint32_t* fown;     // kind << 31 | owner
  const uint32_t* lit_pid;  // literal index -> pattern id
  // Outputs.
  Rec* lit_out;
  uint32_t lit_cap;
  Rec* cand_out;
  uint32_t cand_cap;
  uint32_t* counters;  // [0] matches, [1] candidates, [2] raw regex matches, [3] flags
  uint32_t* queue;     // device work queue head (zeroed per launch), or null: static schedule
  // Profiling: {earliest block start, latest block end} of the prefilter
  // kernel on the device wall clock, or null.
  unsigned long long* stamp;
};

struct HashParams {
  const uint32_t* bloom;
  uint32_t nwords, wshift, wbits, B, NBK, nbits;
  uint32_t s2;  // shift of the second bloom bit (0 = single-bit exact filter)
  uint32_t lane_bytes;  // positions per lane per work unit: 16, 32 or 64
  // Bloom hash for K >= 3 (two 16-bit lanes): multipliers, their squares
  // (windows every second position), -multiplier^K, and the bloom
  // byte-address mask ((nwords - 1) << 2).
  uint32_t BB, BB2, NBKK, amask;
  // Verification table: 2^nbits buckets of one 64-byte entry:
  // {gram key, factor byte offset (16-aligned) | gram offset in the factor,
  // length | kind << 31, output id}, {bytes 0-15}, {bytes 16-31},
  // {first, count, key 0, key 1} of the bucket's further entries in vmore
  // (same entry format; key kVtabEmpty past count). Key kVtabEmpty: an
  // empty bucket.
  const uint4* vtab;
  const uint4* vmore;
  const uint4* fmeta;   // per factor {byte offset (4-aligned), length, kind << 31 | owner, output id}
};
constexpr uint32_t kVtabEmpty = 0xFFFFFFFFu;

// Per-block output slices (no shared counter in the hot path) plus a shared
// overflow area; k_slice_offsets2 + k_slice_gather2 compact them afterwards.
struct SliceOut {
  Rec* buf;
  uint32_t cap;        // records per slice
  uint32_t* counts;    // per slice
  Rec* ovf;
  uint32_t ovf_cap;
  uint32_t* ovf_count;
};

struct AcParams {
  const uint32_t* table32;
  const uint16_t* table16;
  const uint8_t* cmap;
  const uint32_t* out_start;
  const uint32_t* out_list;
  const uint32_t* aclen;  // per factor: length the automaton recognises
  uint32_t states, classes, first_out, maxlen;
};

struct WalkParams {
  const uint8_t* base;
  uint64_t load_end;  // bytes [0, load_end) of base may be loaded
  const uint2* items;
  const Rec* cands;
  const uint32_t* ncand;
  uint32_t cand_cap;
  const uint32_t* tables;
  const uint32_t* toff;
  const uint32_t* ncls;
  const uint8_t* cmaps;
  // Per factor slot (byte offset / 16) and start context: the state after
  // the factor's bytes and where matches ended inside it (a candidate whose
  // z field is slot + 1 starts its walk after the factor); null: no resume.
  const uint2* resume;
  const uint32_t* rx_pid;
  Rec* raw;
  uint32_t raw_cap;
  uint32_t* counters;
  uint32_t window_truncated;  // segment ends before the true input end
  // Per regex: its four start states (preceded by the text start, '\n', a
  // word byte, another byte); the table row stride is ncls (its classes plus
  // the end-of-text column, the last).
  const uint32_t* starts;
  uint32_t ctx_before;  // the byte before a segment's start is in the buffer (a stream chunk after the first)
  // k_rx_scan: regexes whose candidate starts are every position: their
  // indices, and the positions per regex (npos_total, over all segments; a
  // batch lists each item's first position in pos_base, nitems + 1 entries).
  const uint32_t* scan_rx;
  uint32_t nscan;
  uint32_t nitems;
  uint64_t npos_total;
  const uint64_t* pos_base;
  // Optional (device-sized runs): chain ordering pairs written with each
  // raw match, key ((item << rbits | regex) << sbits) | start and value its
  // raw index, counted in bucket key >> cshift.
  uint64_t* ckeys;
  uint32_t* cvals;
  uint32_t* ccnt;
  uint32_t cshift, sbits, rbits;
  // A batch: chain keys (regex << sbits) | (cbase[item] + start), the dense
  // position (a group (item, regex) stays contiguous and the buckets follow
  // positions, not items).
  const uint64_t* cbase;
};
// Chain kernels: groups given by the records' (regex, item) instead of key bits.
constexpr uint32_t kGroupByRecord = 65;

// Batch ordering over the items laid end to end in id order: item i's start
// s is position bases[i] + s (bases: exclusive prefix sums of item length + 1,
// nitems + 1 entries), so ordering keys (position << pbits | pattern id) are
// dense, and tagged output finds the item again by binary search. bases null:
// one segment, position (item << sbits) | start as before.
struct DenseOut {
  const uint64_t* bases;
  const uint64_t* ids;  // source id per item (tagged output)
  uint32_t nitems;
};

__device__ __forceinline__ uint64_t order_pos(const DenseOut& d, uint32_t item, uint32_t start, uint32_t sbits) {
  return d.bases ? d.bases[item] + start : (((uint64_t)item << sbits) | start);
}

// Optional output of k_chain_walk: final ordering pairs of the kept matches,
// key (position << pbits) | pattern id (position as order_pos) and value
// end, counted in bucket key >> shift.
struct OrderOut {
  uint64_t* keys;
  uint32_t* vals;
  uint32_t* cnt;
  uint32_t shift, sbits, pbits;
  DenseOut dense;
};

enum : uint32_t { kFlagTruncatedWalk = 1u };

// Capacities of a device-sized run (no host readback between passes), and
// the counters slot that keeps the literal count before regex matches are
// appended.
// Up to kCallItemsMax run segments passed to k_call_init by value, so a call
// needs no host-to-device copy for them.
constexpr uint32_t kCallItemsMax = 32;
struct CallItems {
  uint32_t n;
  uint2 v[kCallItemsMax];
};

struct RunCaps {
  uint32_t lit_ovf_cap, cand_ovf_cap;
  uint32_t lit_cap, cand_cap, raw_cap, rec_cap, max_matches;
};
constexpr uint32_t kCounterLitSnapshot = 10;
// The counter buffer: the run counters (words 0-15), then two filter
// counter blocks of 16 words (two-stream runs: the filter of run k counts
// its hits and slice overflows in block k % 2, words 5, 6 and 8).
constexpr uint32_t kCounterWords = 48;
constexpr uint32_t kFilterCounterBlock = 16;
// Literal-only ordering: records per bucket of the fixed-capacity buckets,
// and the most buckets (the count array always has this many entries).
constexpr uint32_t kLitBucketCap = 256;
constexpr uint32_t kLitBuckets = 8192;
// Bucket sorts of up to this many buckets scan their counts in one block;
// their count and cursor arrays always hold this many entries (the counts
// beyond the buckets in use zero), so the scan loads them unconditionally.
constexpr uint32_t kScanOneBlock = 1u << 16;
// sticky[0]: a buffer overflowed (redo on the checked path); sticky[1]: error
// bits (1 = regex window truncated, 2 = more than max_matches);
// sticky[2..4]: running sums of filter hits, candidates, raw regex matches.
enum : uint32_t { kStickyRedo = 0, kStickyError = 1, kStickyHits = 2, kStickyCands = 3, kStickyRaw = 4 };
// A call's state block: sticky words, the 64-bit match total at words 8-9
// (both zeroed by k_call_init), kernel stamp pairs from word 16.
constexpr uint32_t kStateResetWords = 10;
// One-block ordering of a literal-only run (k_lit_order_small): at most this
// many matches, from at most kSmallOrderSlices filter blocks.
constexpr uint32_t kSmallOrder = 4096;
constexpr uint32_t kSmallOrderSlices = 1023;

}  // namespace ws
