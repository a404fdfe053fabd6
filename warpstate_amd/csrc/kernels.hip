// HIP kernels of the scan path, written for gfx950 (CDNA4, wave64).
//
//  k_hash_filter rolling-hash prefilter (16-64 positions per lane, bloom
//                filter in LDS) with the verification of its hits in the
//                same kernel: per-wave LDS hit rings, 64 hits a round
//                (filter.hpp; instantiated in filter_k*.hip).
//  k_slice_*     compaction of the per-block output slices.
//  k_ac_scan     dense factor-DFA (Aho-Corasick over byte classes) traversal,
//                one chunk per lane, table in LDS or L2.
//  k_rx_walk     anchored leftmost-first DFA walk from each regex candidate.
//  k_chain_*     per-group greedy non-overlap chains (find_iter semantics).
//  k_*_keys/pack ordering keys and output packing.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "filter.hpp"
#include "kernels.hpp"
#include "regex.hpp"

namespace ws {

// Slice compaction of both verify outputs. One 1024-lane block: lanes
// [0, 512) take an exclusive scan of min(count, cap) over the literal
// slices, lanes [512, 1024) over the candidate slices; offsets[n] is the
// slices' sum, and the total (overflow area included) lands in *lit_total /
// *cand_total, the literal total also in *lit_snap when set.
__global__ __launch_bounds__(1024) void k_slice_offsets2(SliceOut lit, SliceOut cand, uint32_t n, uint32_t* loffs,
                                                          uint32_t* coffs, uint32_t* lit_total, uint32_t* cand_total,
                                                          uint32_t* lit_snap) {
  // Each half's prefix: DPP wave prefixes, then the half's eight wave
  // totals through LDS (one barrier).
  __shared__ uint32_t s_wt[16];
  const uint32_t half = threadIdx.x >> 9, t = threadIdx.x & 511;
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const SliceOut& so = half ? cand : lit;
  uint32_t* offsets = half ? coffs : loffs;
  const uint32_t per = (n + 511) / 512;
  const uint32_t b0 = t * per;
  uint32_t sum = 0;
  for (uint32_t k = 0; k < per && b0 + k < n; ++k) sum += min(so.counts[b0 + k], so.cap);
  const uint32_t incl = wave_incl_sum(sum, lane);
  if (lane == 63) s_wt[w] = incl;
  __syncthreads();
  uint32_t before = 0;
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) before += k < (w & 7u) ? s_wt[(w & 8u) + k] : 0u;
  uint32_t run = before + incl - sum;
  for (uint32_t k = 0; k < per && b0 + k < n; ++k) {
    offsets[b0 + k] = run;
    run += min(so.counts[b0 + k], so.cap);
  }
  if (t == 511) {
    const uint32_t all = before + incl;
    offsets[n] = all;
    const uint32_t tot = all + min(*so.ovf_count, so.ovf_cap);
    if (half) {
      *cand_total = tot;
    } else {
      *lit_total = tot;
      if (lit_snap) *lit_snap = tot;
    }
  }
}

// Gathers both outputs into contiguous lists: blocks [0, n + 64) serve the
// literal slices, the rest the candidate slices; in each half block b < n
// copies slice b and the 64 blocks after them the overflow area.
__global__ __launch_bounds__(kTPB) void k_slice_gather2(SliceOut lit, SliceOut cand, uint32_t n,
                                                         const uint32_t* loffs, const uint32_t* coffs, Rec* lout,
                                                         uint32_t lcap, Rec* cout, uint32_t ccap) {
  const uint32_t span = n + 64;
  const bool second = blockIdx.x >= span;
  const SliceOut& so = second ? cand : lit;
  const uint32_t* offsets = second ? coffs : loffs;
  Rec* out = second ? cout : lout;
  const uint32_t out_cap = second ? ccap : lcap;
  const uint32_t bx = second ? blockIdx.x - span : blockIdx.x;
  if (bx < n) {
    const uint32_t m = min(so.counts[bx], so.cap), o = offsets[bx];
    for (uint32_t k = threadIdx.x; k < m; k += kTPB)
      if (o + k < out_cap) out[o + k] = so.buf[(size_t)bx * so.cap + k];
  } else {
    const uint32_t no = min(*so.ovf_count, so.ovf_cap), o = offsets[n];
    for (uint32_t k = (bx - n) * kTPB + threadIdx.x; k < no; k += 64 * kTPB)
      if (o + k < out_cap) out[o + k] = so.ovf[k];
  }
}

// Literal matches from the slices straight to ordering pairs (and, when
// cout is set, the candidate slices to records as k_slice_gather2's second
// half, on blocks n + 64 on): blocks as the literal half of k_slice_gather2; record
// {pid, start, end, item} becomes key (((item << sbits) | start) << pbits) |
// pid with value end (as k_order_keys makes them), counted in bucket
// key >> shift for the bucket sort that follows.
// Sum over slices [0, upto) of the records they hold, in every thread of
// the block (one call per block).
__device__ __forceinline__ uint32_t block_slice_sum(const SliceOut& so, uint32_t upto) {
  __shared__ uint32_t red[kTPB / 64];
  uint32_t v = 0;
  for (uint32_t b = threadIdx.x; b < upto; b += kTPB) v += min(so.counts[b], so.cap);
  v = __builtin_amdgcn_readlane(wave_incl_sum(v, threadIdx.x & 63), 63);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  uint32_t t = 0;
#pragma unroll
  for (uint32_t w = 0; w < kTPB / 64; ++w) t += red[w];
  return t;
}

// Compaction of the filter's output slices, each block finding its own
// output offset from the slice counts before it: blocks [0, n) copy the
// literal slices straight to ordering pairs (and bucket counts), blocks
// [n, n + 64) the literal overflow area (block n also writes the literal
// total), and with cout set the same again for the candidate slices, as
// records.
__global__ __launch_bounds__(kTPB) void k_slice_gather_keys(SliceOut lit, uint32_t n, uint32_t cap,
                                                             uint32_t sbits, uint32_t pbits, uint64_t* keys,
                                                             uint32_t* vals, uint32_t* cnt, uint32_t shift,
                                                             SliceOut cand, Rec* cout, uint32_t ccap,
                                                             uint32_t* lit_total, uint32_t* cand_total,
                                                             uint32_t* lit_snap, DenseOut dense) {
  if (blockIdx.x >= n + 64) {
    const uint32_t bx = blockIdx.x - (n + 64);
    const uint32_t o = block_slice_sum(cand, bx < n ? bx : n);
    if (bx < n) {
      const uint32_t m = min(cand.counts[bx], cand.cap);
      for (uint32_t k = threadIdx.x; k < m; k += kTPB)
        if (o + k < ccap) cout[o + k] = cand.buf[(size_t)bx * cand.cap + k];
    } else {
      const uint32_t no = min(*cand.ovf_count, cand.ovf_cap);
      if (bx == n && threadIdx.x == 0) *cand_total = o + no;
      for (uint32_t k = (bx - n) * kTPB + threadIdx.x; k < no; k += 64 * kTPB)
        if (o + k < ccap) cout[o + k] = cand.ovf[k];
    }
    return;
  }
  auto put = [&](uint32_t i, const Rec& r) {
    const uint64_t k = (order_pos(dense, r.w, r.y, sbits) << pbits) | r.x;
    keys[i] = k;
    vals[i] = r.z;
    atomicAdd(&cnt[k >> shift], 1u);
  };
  const uint32_t o = block_slice_sum(lit, blockIdx.x < n ? blockIdx.x : n);
  if (blockIdx.x < n) {
    const uint32_t m = min(lit.counts[blockIdx.x], lit.cap);
    for (uint32_t k = threadIdx.x; k < m; k += kTPB)
      if (o + k < cap) put(o + k, lit.buf[(size_t)blockIdx.x * lit.cap + k]);
  } else {
    const uint32_t no = min(*lit.ovf_count, lit.ovf_cap);
    if (blockIdx.x == n && threadIdx.x == 0) {
      *lit_total = o + no;
      if (lit_snap) *lit_snap = o + no;
    }
    for (uint32_t k = (blockIdx.x - n) * kTPB + threadIdx.x; k < no; k += 64 * kTPB)
      if (o + k < cap) put(o + k, lit.ovf[k]);
  }
}

// ---------------------------------------------------------------------------
// Factor-DFA prefilter. The block stages its tile (kTPB * CH bytes plus a
// lookahead) in LDS with 16-byte loads; every lane walks its CH-byte chunk
// from the root state and keeps factors that start inside the chunk.
// ---------------------------------------------------------------------------
template <bool LDS_TABLE, int CH>
__global__ __launch_bounds__(kTPB) void k_ac_scan(ScanCommon c, AcParams a) {
  extern __shared__ uint32_t smem[];
  constexpr uint32_t kTile = kTPB * CH;
  constexpr uint32_t kLook = 16;  // >= maxlen - 1 (maxlen <= 12)
  uint8_t* tile = reinterpret_cast<uint8_t*>(smem);
  uint8_t* cmap = tile + kTile + kLook;
  Rec* s_lit = reinterpret_cast<Rec*>(cmap + 256);
  Rec* s_cand = s_lit + kRecCap;
  uint16_t* table = reinterpret_cast<uint16_t*>(s_cand + kRecCap);
  __shared__ uint32_t s_n[2];
  __shared__ uint32_t s_base[4];
  stamp_start(c);

  for (uint32_t k = threadIdx.x; k < 256; k += kTPB) cmap[k] = a.cmap[k];
  if (LDS_TABLE) {
    const uint32_t n = a.states * a.classes;
    for (uint32_t k = threadIdx.x; k < n; k += kTPB) table[k] = a.table16[k];
  }
  if (threadIdx.x == 0) { s_n[0] = 0; s_n[1] = 0; }
  __syncthreads();

  __shared__ uint32_t s_slot;
  uint32_t ahead = 0;
  for (uint32_t u = next_group(c, 0, true, &s_slot, ahead); u < c.nunits;
       u = next_group(c, u, false, &s_slot, ahead)) {
    uint64_t tile_base;
    uint32_t item;
    if (c.units) {
      const uint4 uu = c.units[u];
      tile_base = uu.x;
      item = uu.y;
    } else {
      tile_base = (uint64_t)u * kTile;
      item = 0;
    }
    const uint2 seg = c.items[item];
    const uint64_t lo = seg.x, hi = seg.y;
    uint64_t lend = (hi + 15) & ~15ull;
    if (lend > c.load_end) lend = c.load_end;
    // Stage [tile_base, tile_base + kTile + kLook) into LDS; only 16-byte
    // blocks that overlap the segment are read.
    for (uint32_t k = threadIdx.x; k < (kTile + kLook) / 16; k += kTPB) {
      uint64_t g = tile_base + 16ull * k;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (g + 16 <= lend && g + 16 > (lo & ~15ull)) v = *reinterpret_cast<const uint4*>(c.base + g);
      *reinterpret_cast<uint4*>(tile + 16 * k) = v;
    }
    __syncthreads();
    const uint64_t cs0 = tile_base + (uint64_t)CH * threadIdx.x;
    const uint64_t cs = cs0 > lo ? cs0 : lo;
    const uint64_t ce = cs0 + CH;
    uint64_t stop = ce + a.maxlen - 1;
    if (stop > hi) stop = hi;
    uint64_t em = lo + c.emit_end;
    uint32_t nl = 0, nc = 0;
    if (cs < ce && cs < hi && cs < em) {
      uint32_t s = 0;
      for (uint64_t e = cs; e < stop; ++e) {
        uint8_t b = tile[e - tile_base];
        uint32_t cls = cmap[b];
        s = LDS_TABLE ? (uint32_t)table[s * a.classes + cls] : a.table32[(size_t)s * a.classes + cls];
        if (s >= a.first_out) {
          const uint32_t o0 = a.out_start[s - a.first_out], o1 = a.out_start[s - a.first_out + 1];
          for (uint32_t o = o0; o < o1; ++o) {
            const uint32_t f = a.out_list[o];
            const uint64_t st = e + 1 - a.aclen[f];
            if (st < cs || st >= ce || st >= em) continue;
            if (!verify_factor(c, st, (uint32_t)hi, f)) continue;
            const uint32_t own = c.fown[f];
            const uint32_t start = (uint32_t)(st - lo);
            if (own >> 31) {
              uint32_t slot = atomicAdd(&s_n[1], 1u);
              stage_rec(make_uint4(own & 0x7FFFFFFFu, start, 0u, item), slot, s_cand,
                        &c.counters[1], c.cand_out, c.cand_cap);
              ++nc;
            } else {
              uint32_t slot = atomicAdd(&s_n[0], 1u);
              stage_rec(make_uint4(c.lit_pid[own], start, start + c.flen[f], item), slot, s_lit,
                        &c.counters[0], c.lit_out, c.lit_cap);
              ++nl;
            }
          }
        }
      }
    }
    (void)nl;
    (void)nc;
    __syncthreads();
    if (s_n[0] > kRecCap / 2 || s_n[1] > kRecCap / 2) flush_block(s_n, s_lit, s_cand, s_base, c);
  }
  flush_block(s_n, s_lit, s_cand, s_base, c);
  stamp_end(c);
}

// Block-wide reservation of `want` slots per thread from *counter: every
// thread of the block calls it (block-uniform control flow) and gets the
// first of its slots. One device atomic per block: same-address device
// atomics serialise, and one per wave costs microseconds per launch.
// s_res: two LDS words of the caller's, reused call after call (the reset
// of the total follows the last barrier every thread has passed).
__device__ __forceinline__ uint32_t block_reserve(uint32_t want, uint32_t* counter, uint32_t* s_res) {
  // A wave prefix sum (DPP) and one LDS atomic per wave, then one device
  // atomic per block.
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t incl = wave_incl_sum(want, lane);
  const uint32_t wtot = __builtin_amdgcn_readlane(incl, 63);
  if (threadIdx.x == 0) s_res[0] = 0;
  __syncthreads();
  uint32_t wbase = 0;
  if (lane == 63 && wtot) wbase = atomicAdd(&s_res[0], wtot);
  wbase = __builtin_amdgcn_readlane(wbase, 63);
  __syncthreads();
  if (threadIdx.x == 0) s_res[1] = s_res[0] ? atomicAdd(counter, s_res[0]) : 0u;
  __syncthreads();
  return s_res[1] + wbase + incl - want;
}

// ---------------------------------------------------------------------------
// Anchored leftmost-first walks.
// ---------------------------------------------------------------------------
// The context a byte leaves for the position after it: 1 '\n', 2 a word
// byte [0-9A-Za-z_], 3 another byte (0 is the text start).
__device__ __forceinline__ uint32_t byte_ctx(uint32_t b) {
  const uint32_t lo = b | 0x20u;
  const bool w = (b - '0' < 10u) || (lo - 'a' < 26u) || b == '_';
  return b == '\n' ? 1u : (w ? 2u : 3u);
}

// The leftmost-first match of regex rx anchored at st (relative to segment
// seg): its end (relative), or 0xFFFFFFFF. The walk starts in the state for
// the byte before st (the text start at a segment's start, unless the buffer
// holds the byte before it), records every position whose transition says a
// match ends there, and stops when the DFA dies. At the segment's end it
// reads the end-of-text column, or, in a window that ends before the input
// does, sets *open (the match could still change).
// The 16 bytes at a (16-aligned) of the walk's buffer; bytes at or past
// load_end read as zero.
__device__ __forceinline__ uint4 load16(const WalkParams& w, uint64_t a) {
  if (a + 16 <= w.load_end) return *reinterpret_cast<const uint4*>(w.base + a);
  uint32_t v[4] = {0, 0, 0, 0};
  for (uint32_t k = 0; k < 16 && a + k < w.load_end; ++k) v[k >> 2] |= (uint32_t)w.base[a + k] << (8 * (k & 3));
  return make_uint4(v[0], v[1], v[2], v[3]);
}

// fz: the candidate's factor slot + 1 (0: none), see WalkParams::resume.
__device__ __forceinline__ uint32_t rx_match_at(const WalkParams& w, uint32_t rx, uint2 seg, uint32_t st,
                                                bool* open, uint32_t fz = 0) {
  const uint32_t* T = w.tables + w.toff[rx];
  const uint32_t C = w.ncls[rx];
  const uint8_t* cm = w.cmaps + 256u * rx;
  const uint64_t p0 = (uint64_t)seg.x + st;
  const uint32_t ctx = (st > 0 || w.ctx_before) ? byte_ctx(w.base[p0 - 1]) : 0u;
  uint32_t s = w.starts[4 * rx + ctx];
  uint32_t last = 0xFFFFFFFFu;
  uint64_t pb = p0;  // first byte the walk reads
  if (fz && w.resume) {
    const uint2 r = w.resume[4 * (fz - 1) + ctx];
    if (r.y & 0xFFFFu) last = st + (r.y & 0xFFFFu) - 1u;
    if (r.x & 0x80000000u) return last;
    s = r.x;
    pb = p0 + (r.y >> 16);
  }
  // The first 48 bytes from p0 & ~15 in three loads issued together, and
  // their classes looked up at once (independent reads, packed one per
  // byte): inside this window the walk waits on one dependent read per
  // byte, the transition. (A wave's lanes step through their bytes
  // together but cross 16-byte boundaries at different steps, so per-block
  // loads would queue one round trip per crossing.) Past the window, byte by
  // byte.
  const uint64_t a0 = p0 & ~15ull;
  uint32_t cls[12];
  {
    const uint4 b0 = load16(w, a0), b1 = load16(w, a0 + 16), b2 = load16(w, a0 + 32);
    const uint32_t d[12] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w, b2.x, b2.y, b2.z, b2.w};
#pragma unroll
    for (int k = 0; k < 12; ++k)
      cls[k] = (uint32_t)cm[d[k] & 0xFFu] | (uint32_t)cm[(d[k] >> 8) & 0xFFu] << 8 |
               (uint32_t)cm[(d[k] >> 16) & 0xFFu] << 16 | (uint32_t)cm[d[k] >> 24] << 24;
  }
  for (uint64_t p = pb; p < seg.y; ++p) {
    const uint64_t o = p - a0;
    uint32_t c;
    if (o < 48) {
      // cls[o / 4] by selects on the index bits (no indexed register array).
      const uint32_t i = (uint32_t)o >> 2;
      const uint32_t l0 = (i & 1) ? cls[1] : cls[0], l1 = (i & 1) ? cls[3] : cls[2];
      const uint32_t l2 = (i & 1) ? cls[5] : cls[4], l3 = (i & 1) ? cls[7] : cls[6];
      const uint32_t l4 = (i & 1) ? cls[9] : cls[8], l5 = (i & 1) ? cls[11] : cls[10];
      const uint32_t m0 = (i & 2) ? l1 : l0, m1 = (i & 2) ? l3 : l2, m2 = (i & 2) ? l5 : l4;
      const uint32_t v = (i & 8) ? m2 : ((i & 4) ? m1 : m0);
      c = (v >> (8 * ((uint32_t)o & 3))) & 0xFFu;
    } else {
      c = cm[w.base[p]];
    }
    const uint32_t e = T[s * C + c];
    if (e & DFA_MATCH) last = (uint32_t)(p - seg.x);
    if (e & DFA_DEAD) return last;
    s = e & DFA_STATE_MASK;
  }
  if (w.window_truncated) {
    *open = true;
    return last;
  }
  if (T[s * C + C - 1] & DFA_MATCH) last = seg.y - seg.x;
  return last;
}

// Appends a raw match {rx, start, end, item} (block-wide slot reservation;
// every lane of the block calls it) and, in device-sized runs, its chain
// ordering pair.
__device__ __forceinline__ void emit_raw(const WalkParams& w, bool found, uint32_t rx, uint32_t st, uint32_t end,
                                         uint32_t item, uint32_t* s_res) {
  const uint32_t slot = block_reserve(found ? 1u : 0u, &w.counters[2], s_res);
  if (found && slot < w.raw_cap) {
    w.raw[slot] = make_uint4(rx, st, end, item);
    if (w.ckeys) {
      const uint64_t k = w.cbase ? (((uint64_t)rx << w.sbits) | (w.cbase[item] + st))
                                 : ((((uint64_t)item << w.rbits | rx) << w.sbits) | st);
      w.ckeys[slot] = k;
      w.cvals[slot] = slot;
      atomicAdd(&w.ccnt[k >> w.cshift], 1u);
    }
  }
}

// One lane per candidate start (a prefilter hit of a regex factor).
__global__ __launch_bounds__(kTPB) void k_rx_walk(WalkParams w) {
  __shared__ uint32_t s_res[2];
  const uint32_t n = min(*w.ncand, w.cand_cap);
  // Block-uniform loop: every lane reaches the slot reservation.
  for (uint32_t b0 = blockIdx.x * kTPB; b0 < n; b0 += gridDim.x * kTPB) {
    const uint32_t i = b0 + threadIdx.x;
    Rec cd = make_uint4(0, 0, 0, 0);
    uint32_t last = 0xFFFFFFFFu;
    if (i < n) {
      cd = w.cands[i];
      bool open = false;
      last = rx_match_at(w, cd.x, w.items[cd.w], cd.y, &open, cd.z);
      if (open) atomicOr(&w.counters[3], kFlagTruncatedWalk);
    }
    emit_raw(w, last != 0xFFFFFFFFu, cd.x, cd.y, last, cd.w, s_res);
  }
}

// The same walk straight from the filter's candidate slices (device-sized
// runs; no compaction pass): block b < nslices walks slice b, blocks
// [nslices, nslices + kOvfBlocks) the overflow area, and block nslices also
// writes the candidate total (the slice counts' sum plus the overflow's).
constexpr uint32_t kOvfBlocks = 64;

__global__ __launch_bounds__(kTPB) void k_rx_walk_slices(WalkParams w, SliceOut cand, uint32_t nslices,
                                                         uint32_t* cand_total) {
  __shared__ uint32_t s_res[2];
  const bool ovf = blockIdx.x >= nslices;
  const uint32_t m = ovf ? min(*cand.ovf_count, cand.ovf_cap) : min(cand.counts[blockIdx.x], cand.cap);
  const Rec* src = ovf ? cand.ovf : cand.buf + (size_t)blockIdx.x * cand.cap;
  const uint32_t first = ovf ? (blockIdx.x - nslices) * kTPB : 0u;
  const uint32_t step = ovf ? kOvfBlocks * kTPB : kTPB;
  if (blockIdx.x == nslices) {
    const uint32_t t = block_slice_sum(cand, nslices);
    if (threadIdx.x == 0) *cand_total = t + m;
  }
  // Block-uniform loop: every lane reaches the slot reservation.
  for (uint32_t b0 = first; b0 < m; b0 += step) {
    const uint32_t i = b0 + threadIdx.x;
    Rec cd = make_uint4(0, 0, 0, 0);
    uint32_t last = 0xFFFFFFFFu;
    if (i < m) {
      cd = src[i];
      bool open = false;
      last = rx_match_at(w, cd.x, w.items[cd.w], cd.y, &open, cd.z);
      if (open) atomicOr(&w.counters[3], kFlagTruncatedWalk);
    }
    emit_raw(w, last != 0xFFFFFFFFu, cd.x, cd.y, last, cd.w, s_res);
  }
}

// Every position of every segment, for each regex without a selective
// prefix (w.scan_rx): position q of the flat position space lies in item i
// when pos_base[i] <= q < pos_base[i + 1] (a single segment: item 0). An
// item of length L has L + 1 positions (an empty match may start at its
// end). Lanes take consecutive positions, so the first bytes load coalesced.
__global__ __launch_bounds__(kTPB) void k_rx_scan(WalkParams w) {
  __shared__ uint32_t s_res[2];
  const uint64_t total = w.npos_total * w.nscan;
  for (uint64_t b0 = (uint64_t)blockIdx.x * kTPB; b0 < total; b0 += (uint64_t)gridDim.x * kTPB) {
    const uint64_t g = b0 + threadIdx.x;
    uint32_t last = 0xFFFFFFFFu, rx = 0, item = 0, st = 0;
    if (g < total) {
      const uint32_t k = (uint32_t)(g / w.npos_total);
      const uint64_t q = g - (uint64_t)k * w.npos_total;
      rx = w.scan_rx[k];
      uint64_t q0 = 0;
      if (w.pos_base) {
        uint32_t lo = 0, hi = w.nitems;  // pos_base[lo] <= q < pos_base[hi]
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) / 2;
          if (w.pos_base[mid] <= q) lo = mid;
          else hi = mid;
        }
        item = lo;
        q0 = w.pos_base[lo];
      }
      st = (uint32_t)(q - q0);
      bool open = false;
      last = rx_match_at(w, rx, w.items[item], st, &open);
      if (open) atomicOr(&w.counters[3], kFlagTruncatedWalk);
    }
    emit_raw(w, last != 0xFFFFFFFFu, rx, st, last, item, s_res);
  }
}

// Element count of a pass: the host's n, or (device-sized runs) the count
// a previous kernel left in device memory, capped at n.
__device__ __forceinline__ uint32_t count_of(uint32_t n, const uint32_t* nd) { return nd ? min(*nd, n) : n; }

// Chain keys: group (item, regex) in the high bits, start in the low bits;
// with dbase (a batch) regex, then the dense position dbase[item] + start.
__global__ void k_chain_keys(const Rec* raw, uint32_t n, const uint32_t* nd, uint32_t sbits, uint32_t rbits,
                             uint64_t* keys, uint32_t* vals, const uint64_t* dbase) {
  n = count_of(n, nd);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const Rec r = raw[i];
    keys[i] = dbase ? (((uint64_t)r.x << sbits) | (dbase[r.w] + r.y)) : ((((uint64_t)r.w << rbits | r.x) << sbits) | r.y);
    vals[i] = i;
  }
}

// A record starts a chain segment when it opens its group or when no earlier
// record of the group can overlap it (none within max length ends after it).
// group = key >> gshift (all one group when gshift is 64), or with
// kGroupByRecord the record's (regex, item); max length comes from
// maxlen[rec.x] or, when maxlen is null, from maxlen_all.
__device__ __forceinline__ uint64_t group_of(uint64_t k, uint32_t gshift) {
  return gshift >= 64 ? 0ull : (k >> gshift);
}
__device__ __forceinline__ uint64_t group_at(const uint64_t* keys, const uint32_t* vals, const Rec* raw, uint32_t j,
                                             uint32_t gshift) {
  if (gshift == kGroupByRecord) {
    const Rec r = raw[vals[j]];
    return ((uint64_t)r.x << 32) | r.w;
  }
  return group_of(keys[j], gshift);
}

// Whether record i, inside its group past the group's first record, starts
// a chain segment of its own: no earlier record of the group within the
// regex's maximum length ends after its start (an empty record also
// depends on an earlier end equal to its start: find_iter skips an empty
// match where the previous one ended). Unbounded regexes: never.
__device__ __forceinline__ bool chain_starts_inside(const uint64_t* keys, const uint32_t* vals, const Rec* raw,
                                                    uint32_t i, uint64_t g, uint32_t gshift,
                                                    const uint32_t* maxlen, uint32_t maxlen_all) {
  const Rec ri = raw[vals[i]];
  const bool empty = ri.y == ri.z;
  const uint32_t ml = maxlen ? maxlen[ri.x] : maxlen_all;
  if (ml == 0xFFFFFFFFu) return false;
  for (uint32_t j = i; j-- > 0;) {
    if (group_at(keys, vals, raw, j, gshift) != g) break;
    const Rec rj = raw[vals[j]];
    const uint64_t reach = (uint64_t)rj.y + ml;
    if (reach < ri.y || (reach == ri.y && !empty)) break;
    if (rj.z > ri.y || (empty && rj.z == ri.y)) return false;
  }
  return true;
}

// Greedy non-overlapping selection inside each segment (find_iter semantics):
// keep a record when it starts at or after the end of the last kept one.
// With `ties`, records sharing the last kept (start, end) are kept too
// (HashScanner reports every id of the shortest length at a position).
// Output pattern id is rx_pid[rec.x] when rx_pid is set, else rec.x.
// Each lane tests whether its record starts a chain segment (the first of
// its group, or chain_starts_inside); a segment's first lane walks it up
// to the next segment start, twice: once to count the records it keeps,
// then, after the block has reserved all its slots at once, to write them.
// Stream chunks: carried[rx] is 1 + the absolute end of the regex's last
// kept match in earlier chunks (0: none), read-only in this launch;
// carried_out, the other half of a double buffer, is raised to carried
// and to this launch's kept ends (atomicMax in any order: both halves
// start at zero and only grow, so the half written two chunks ago is at
// most carried). Both passes of a lane therefore see the same carried
// value.
// Launched with kTPB threads per block.
__global__ __launch_bounds__(kTPB) void k_chain_walk(const uint64_t* keys, const uint32_t* vals, const Rec* raw,
                                                     uint32_t n, const uint32_t* nd, uint32_t gshift,
                                                     const uint32_t* maxlen, uint32_t maxlen_all,
                                                     const uint32_t* carried, uint32_t* carried_out,
                                                     uint32_t ncarry, uint32_t seg_off, const uint32_t* rx_pid,
                                                     uint32_t ties, Rec* out, uint32_t out_cap,
                                                     uint32_t* out_count, OrderOut oo) {
  __shared__ uint32_t s_res[2];
  if (carried_out)
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < ncarry; k += gridDim.x * blockDim.x)
      atomicMax(&carried_out[k], carried[k]);
  n = count_of(n, nd);
  // Writes kept record r to output slot sl (and its ordering pair, and the
  // regex's carried end).
  auto emit_rec = [&](const Rec& r, uint32_t sl) {
    const uint32_t pid = rx_pid ? rx_pid[r.x] : r.x;
    if (sl < out_cap) {
      out[sl] = make_uint4(pid, r.y, r.z, r.w);
      if (oo.keys) {
        const uint64_t k = (order_pos(oo.dense, r.w, r.y, oo.sbits) << oo.pbits) | pid;
        oo.keys[sl] = k;
        oo.vals[sl] = r.z;
        atomicAdd(&oo.cnt[k >> oo.shift], 1u);
      }
    }
    if (carried_out) atomicMax(&carried_out[r.x], seg_off + r.z + 1u);
  };
  // The chain from segment start i (group g): with `emit`, its kept records
  // go to slots slot, slot + 1, ...; returns how many it keeps, and
  // (counting) the last kept record in *kept_rec.
  auto walk = [&](uint32_t i, uint64_t g, bool emit, uint32_t slot, Rec* kept_rec) -> uint32_t {
    uint32_t last_end = 0, last_start = 0xFFFFFFFFu, kept = 0;
    bool have_last = false;
    if (carried) {
      const uint32_t cv = carried[raw[vals[i]].x];
      if (cv > seg_off) {  // the previous chunks' last kept match ends at or after this chunk's start
        last_end = cv - 1u - seg_off;
        have_last = true;
      }
    }
    for (uint32_t j = i; j < n; ++j) {
      if (j > i && (group_at(keys, vals, raw, j, gshift) != g ||
                    chain_starts_inside(keys, vals, raw, j, g, gshift, maxlen, maxlen_all)))
        break;
      const Rec r = raw[vals[j]];
      const bool tie = ties && r.y == last_start && r.z == last_end;
      if (r.y < last_end && !tie) continue;
      // find_iter: an empty match where the previous match ended is skipped.
      if (r.y == r.z && have_last && r.y == last_end && !tie) continue;
      last_end = r.z;
      last_start = r.y;
      have_last = true;
      if (emit) emit_rec(r, slot + kept);
      else *kept_rec = r;
      ++kept;
    }
    return kept;
  };
  // Block-uniform loop: every lane reaches the slot reservation. A chain
  // that keeps one record (most) emits it from registers; longer ones are
  // walked again.
  for (uint32_t b0 = blockIdx.x * kTPB; b0 < n; b0 += gridDim.x * kTPB) {
    const uint32_t i = b0 + threadIdx.x;
    uint64_t g = 0;
    bool on = false;
    if (i < n) {
      g = group_at(keys, vals, raw, i, gshift);
      on = i == 0 || group_at(keys, vals, raw, i - 1, gshift) != g ||
           chain_starts_inside(keys, vals, raw, i, g, gshift, maxlen, maxlen_all);
    }
    Rec one = make_uint4(0, 0, 0, 0);
    const uint32_t kept = on ? walk(i, g, false, 0, &one) : 0u;
    const uint32_t slot = block_reserve(kept, out_count, s_res);
    if (kept == 1)
      emit_rec(one, slot);
    else if (kept)
      walk(i, g, true, slot, &one);
  }
}

// Literal reduction keys: (start, length or maxlen-length, pattern_id).
__global__ void k_lit_keys(const Rec* recs, uint32_t n, uint32_t lbits, uint32_t pbits,
                           uint32_t maxlen, uint32_t len_desc, uint64_t* keys, uint32_t* vals) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const Rec r = recs[i];
    const uint32_t len = r.z - r.y;
    const uint64_t lk = len_desc ? (uint64_t)(maxlen - len) : (uint64_t)len;
    keys[i] = (((uint64_t)r.y << lbits | lk) << pbits) | r.x;
    vals[i] = i;
  }
}

// Per-pattern keys: (pattern_id, start).
__global__ void k_pid_keys(const Rec* recs, uint32_t n, uint32_t sbits, uint64_t* keys,
                           uint32_t* vals) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const Rec r = recs[i];
    keys[i] = ((uint64_t)r.x << sbits) | r.y;
    vals[i] = i;
  }
}

// Final ordering keys: (position, pattern_id), the position (item, start)
// or a dense batch position (order_pos); the end rides as value.
__global__ void k_order_keys(const Rec* recs, uint32_t n, const uint32_t* nd, uint32_t sbits, uint32_t pbits,
                             uint64_t* keys, uint32_t* vals, DenseOut dense) {
  n = count_of(n, nd);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const Rec r = recs[i];
    keys[i] = (order_pos(dense, r.w, r.y, sbits) << pbits) | r.x;
    vals[i] = r.z;
  }
}

struct MatchOut { uint32_t pid, start, end; };
struct TaggedOut { uint64_t id; uint32_t pid, start, end, pad; };

// Writes output record `slot` of a match whose ordering key position is pos
// and whose end is `end` (relative to its segment), or, with `len`, whose
// length is `end`: a Match at start_off + pos, or with dense batch positions
// a TaggedMatch (the item is the last one whose base is <= pos).
__device__ __forceinline__ void put_match(void* out, uint64_t slot, uint64_t pos, uint32_t pid, uint32_t end,
                                          uint32_t start_off, const DenseOut& d, bool len = false) {
  if (!d.bases) {
    MatchOut m;
    m.pid = pid;
    m.start = (uint32_t)pos + start_off;
    m.end = len ? m.start + end : end + start_off;
    reinterpret_cast<MatchOut*>(out)[slot] = m;
    return;
  }
  uint32_t lo = 0, hi = d.nitems;  // bases[lo] <= pos < bases[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (d.bases[mid] <= pos) lo = mid;
    else hi = mid;
  }
  TaggedOut t;
  t.id = d.ids[lo];
  t.pid = pid;
  t.start = (uint32_t)(pos - d.bases[lo]);
  t.end = len ? t.start + end : end;
  t.pad = 0;
  reinterpret_cast<TaggedOut*>(out)[slot] = t;
}


// Writes out[base + i] for i < n (base = *base_dev for device-sized runs,
// else 0), dropping records at or past out_cap: Matches, or TaggedMatches
// from dense batch positions.
__global__ void k_pack_matches(const uint64_t* keys, const uint32_t* vals, uint32_t n, const uint32_t* nd,
                               uint32_t pbits, uint32_t start_off, void* out, uint64_t out_cap,
                               const unsigned long long* base_dev, DenseOut dense) {
  n = count_of(n, nd);
  const uint64_t base = base_dev ? *base_dev : 0ull;
  const uint64_t pmask = (1ull << pbits) - 1;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (base + i >= out_cap) break;
    const uint64_t k = keys[i];
    put_match(out, base + i, k >> pbits, (uint32_t)(k & pmask), vals[i], start_off, dense);
  }
}

__global__ void k_pack_tagged(const uint64_t* keys, const uint32_t* vals, uint32_t n,
                              uint32_t sbits, uint32_t pbits, const uint64_t* ids,
                              TaggedOut* out) {
  const uint64_t pmask = (1ull << pbits) - 1, smask = (1ull << sbits) - 1;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint64_t k = keys[i];
    TaggedOut t;
    t.pid = (uint32_t)(k & pmask);
    t.start = (uint32_t)((k >> pbits) & smask);
    t.id = ids[(uint32_t)(k >> (pbits + sbits))];
    t.end = vals[i];
    t.pad = 0;
    out[i] = t;
  }
}

// Batch ordering keys over the items laid end to end in id order: item i's
// match at start s sits at bases[i] + s (bases: exclusive prefix sums of the
// item lengths plus one, the end position of each item), so keys
// (position << pbits | pid) are dense and the bucket sort applies. Same
// order as (item, start, pid).
__global__ void k_order_keys_dense(const Rec* recs, uint32_t n, uint32_t pbits, const uint64_t* bases,
                                   uint64_t* keys, uint32_t* vals) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const Rec r = recs[i];
    keys[i] = ((bases[r.w] + r.y) << pbits) | r.x;
    vals[i] = r.z;
  }
}

// Tagged output from dense batch keys: the item is the last one whose base is
// <= the position (bases increase strictly), found by binary search.
__global__ void k_pack_tagged_dense(const uint64_t* keys, const uint32_t* vals, uint32_t n, uint32_t pbits,
                                    const uint64_t* bases, uint32_t nitems, const uint64_t* ids, TaggedOut* out) {
  const uint64_t pmask = (1ull << pbits) - 1;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint64_t k = keys[i];
    const uint64_t g = k >> pbits;
    uint32_t lo = 0, hi = nitems;  // bases[lo] <= g < bases[hi] (hi = nitems: past the end)
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (bases[mid] <= g) lo = mid;
      else hi = mid;
    }
    TaggedOut t;
    t.pid = (uint32_t)(k & pmask);
    t.start = (uint32_t)(g - bases[lo]);
    t.id = ids[lo];
    t.end = vals[i];
    t.pad = 0;
    out[i] = t;
  }
}

hipError_t launch_order_keys_dense(const Rec* recs, uint32_t n, uint32_t pbits, const uint64_t* bases,
                                   uint64_t* keys, uint32_t* vals, hipStream_t st) {
  const uint32_t g = std::max(1u, std::min(4096u, (n + kTPB - 1) / kTPB));
  hipLaunchKernelGGL(k_order_keys_dense, dim3(g), dim3(kTPB), 0, st, recs, n, pbits, bases, keys, vals);
  return hipGetLastError();
}

hipError_t launch_pack_tagged_dense(const uint64_t* keys, const uint32_t* vals, uint32_t n, uint32_t pbits,
                                    const uint64_t* bases, uint32_t nitems, const uint64_t* ids, void* out,
                                    hipStream_t st) {
  const uint32_t g = std::max(1u, std::min(4096u, (n + kTPB - 1) / kTPB));
  hipLaunchKernelGGL(k_pack_tagged_dense, dim3(g), dim3(kTPB), 0, st, keys, vals, n, pbits, bases, nitems, ids,
                     reinterpret_cast<TaggedOut*>(out));
  return hipGetLastError();
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host launch wrappers (the runtime in engine.cpp only calls these).
// ---------------------------------------------------------------------------
static std::mutex g_lds_mu;
static std::map<const void*, size_t> g_lds_set;

hipError_t allow_lds_bytes(const void* key, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  std::lock_guard<std::mutex> lk(g_lds_mu);
  auto it = g_lds_set.find(key);
  if (it != g_lds_set.end() && it->second >= bytes) return hipSuccess;
  hipError_t e = hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) g_lds_set[key] = bytes;
  return e;
}

// Sets the filter kernel's dynamic LDS limit ahead of launches that may be
// captured into a graph.
hipError_t prepare_hash(uint32_t K, bool ci, uint32_t lane_bytes, uint32_t nwords, bool rx) {
  const size_t shmem = hash_shmem_bytes(nwords);
  switch (K) {
#define WS_K(k) \
  case k: return prepare_hash_k<k>(ci, lane_bytes, shmem, rx);
    WS_K(1) WS_K(2) WS_K(3) WS_K(4) WS_K(5) WS_K(6) WS_K(7) WS_K(8) WS_K(11) WS_K(12) WS_K(15) WS_K(16)
#undef WS_K
    default: return hipErrorInvalidValue;
  }
}

// The filter-and-verify kernel on `grid` blocks; block b writes output
// slice b of lit and cand (rx: the set has regex factors).
hipError_t launch_hash(uint32_t K, bool ci, uint32_t grid, hipStream_t st, const ScanCommon& c,
                       const HashParams& h, const SliceOut& lit, const SliceOut& cand, bool rx) {
  switch (K) {
#define WS_K(k) \
  case k: return launch_hash_k<k>(ci, grid, st, c, h, lit, cand, rx);
    WS_K(1) WS_K(2) WS_K(3) WS_K(4) WS_K(5) WS_K(6) WS_K(7) WS_K(8) WS_K(11) WS_K(12) WS_K(15) WS_K(16)
#undef WS_K
    default: return hipErrorInvalidValue;
  }
}

constexpr int kAcChunk = 64;
uint32_t ac_unit_bytes() { return kTPB * kAcChunk; }
size_t ac_shmem_bytes(bool lds_table, uint32_t states, uint32_t classes) {
  size_t b = kTPB * kAcChunk + 16 + 256 + 2 * kRecCap * sizeof(Rec);
  if (lds_table) b += (size_t)states * classes * 2;
  return b;
}

hipError_t prepare_ac(bool lds_table, uint32_t states, uint32_t classes) {
  const size_t shmem = ac_shmem_bytes(lds_table, states, classes);
  return lds_table ? allow_lds(k_ac_scan<true, kAcChunk>, shmem) : allow_lds(k_ac_scan<false, kAcChunk>, shmem);
}

hipError_t launch_ac(bool lds_table, uint32_t grid, hipStream_t st, const ScanCommon& c,
                     const AcParams& a) {
  const size_t shmem = ac_shmem_bytes(lds_table, a.states, a.classes);
  if (lds_table) {
    hipError_t e = allow_lds(k_ac_scan<true, kAcChunk>, shmem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_ac_scan<true, kAcChunk>), dim3(grid), dim3(kTPB), shmem, st, c, a);
  } else {
    hipError_t e = allow_lds(k_ac_scan<false, kAcChunk>, shmem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_ac_scan<false, kAcChunk>), dim3(grid), dim3(kTPB), shmem, st, c, a);
  }
  return hipGetLastError();
}

// Grid of a pass whose element count lives on the device (grid-stride).
constexpr uint32_t kDeviceGrid = 2048;

static uint32_t grid_for(uint32_t n) {
  uint32_t g = (n + kTPB - 1) / kTPB;
  return g < 1 ? 1 : (g > 65535 ? 65535 : g);
}

// ncand_host sizes the grid; 0 = the count is only known on the device
// (k_rx_walk reads it), so launch a fixed grid-stride grid.
hipError_t launch_walk(uint32_t ncand_host, hipStream_t st, const WalkParams& w) {
  hipLaunchKernelGGL(k_rx_walk, dim3(ncand_host ? grid_for(ncand_host) : kDeviceGrid), dim3(kTPB), 0, st, w);
  return hipGetLastError();
}

hipError_t launch_walk_slices(hipStream_t st, const WalkParams& w, const SliceOut& cand, uint32_t nslices,
                              uint32_t* cand_total) {
  hipLaunchKernelGGL(k_rx_walk_slices, dim3(nslices + kOvfBlocks), dim3(kTPB), 0, st, w, cand, nslices, cand_total);
  return hipGetLastError();
}

// Every position for the regexes of w.scan_rx (grid-stride over
// npos_total * nscan lanes).
hipError_t launch_rx_scan(hipStream_t st, const WalkParams& w) {
  const uint64_t total = w.npos_total * w.nscan;
  const uint64_t g = std::min<uint64_t>(4096, std::max<uint64_t>(1, (total + kTPB - 1) / kTPB));
  hipLaunchKernelGGL(k_rx_scan, dim3((uint32_t)g), dim3(kTPB), 0, st, w);
  return hipGetLastError();
}

// The launchers below take a host count n; with nd set, n is a capacity and
// the kernel reads the actual count from *nd.
static uint32_t grid_of(uint32_t n, const uint32_t* nd) { return nd ? std::min(kDeviceGrid, grid_for(n)) : grid_for(n); }

hipError_t launch_order_keys(const Rec* recs, uint32_t n, const uint32_t* nd, uint32_t sbits, uint32_t pbits,
                             uint64_t* keys, uint32_t* vals, hipStream_t st, const DenseOut& dense) {
  hipLaunchKernelGGL(k_order_keys, dim3(grid_of(n, nd)), dim3(kTPB), 0, st, recs, n, nd, sbits, pbits, keys, vals,
                     dense);
  return hipGetLastError();
}

hipError_t launch_chain_keys(const Rec* raw, uint32_t n, const uint32_t* nd, uint32_t sbits, uint32_t rbits,
                             uint64_t* keys, uint32_t* vals, hipStream_t st, const uint64_t* dbase) {
  hipLaunchKernelGGL(k_chain_keys, dim3(grid_of(n, nd)), dim3(kTPB), 0, st, raw, n, nd, sbits, rbits, keys, vals,
                     dbase);
  return hipGetLastError();
}

hipError_t launch_chain_walk(const uint64_t* keys, const uint32_t* vals, const Rec* raw, uint32_t n,
                             const uint32_t* nd, uint32_t gshift, const uint32_t* maxlen, uint32_t maxlen_all,
                             const uint32_t* carried, uint32_t* carried_out, uint32_t ncarry, uint32_t seg_off,
                             const uint32_t* rx_pid, uint32_t ties, Rec* out, uint32_t out_cap,
                             uint32_t* out_count, hipStream_t st, const OrderOut* oo) {
  hipLaunchKernelGGL(k_chain_walk, dim3(grid_of(n, nd)), dim3(kTPB), 0, st, keys, vals, raw, n, nd, gshift, maxlen,
                     maxlen_all, carried, carried_out, ncarry, seg_off, rx_pid, ties, out, out_cap, out_count,
                     oo ? *oo : OrderOut{});
  return hipGetLastError();
}

hipError_t launch_lit_keys(const Rec* recs, uint32_t n, uint32_t lbits, uint32_t pbits, uint32_t maxlen,
                           uint32_t len_desc, uint64_t* keys, uint32_t* vals, hipStream_t st) {
  hipLaunchKernelGGL(k_lit_keys, dim3(grid_for(n)), dim3(kTPB), 0, st, recs, n, lbits, pbits, maxlen,
                     len_desc, keys, vals);
  return hipGetLastError();
}

hipError_t launch_pid_keys(const Rec* recs, uint32_t n, uint32_t sbits, uint64_t* keys, uint32_t* vals,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_pid_keys, dim3(grid_for(n)), dim3(kTPB), 0, st, recs, n, sbits, keys, vals);
  return hipGetLastError();
}

hipError_t launch_pack_matches(const uint64_t* keys, const uint32_t* vals, uint32_t n, const uint32_t* nd,
                               uint32_t pbits, uint32_t start_off, void* out, uint64_t out_cap,
                               const unsigned long long* base_dev, hipStream_t st, const DenseOut& dense) {
  hipLaunchKernelGGL(k_pack_matches, dim3(grid_of(n, nd)), dim3(kTPB), 0, st, keys, vals, n, nd, pbits, start_off,
                     out, out_cap, base_dev, dense);
  return hipGetLastError();
}

hipError_t launch_pack_tagged(const uint64_t* keys, const uint32_t* vals, uint32_t n, uint32_t sbits,
                              uint32_t pbits, const uint64_t* ids, void* out, hipStream_t st) {
  hipLaunchKernelGGL(k_pack_tagged, dim3(grid_for(n)), dim3(kTPB), 0, st, keys, vals, n, sbits, pbits, ids,
                     reinterpret_cast<TaggedOut*>(out));
  return hipGetLastError();
}


// Compacts the literal and candidate slices (one scan kernel, one gather).
hipError_t launch_slice_compact2(const SliceOut& lit, const SliceOut& cand, uint32_t n, uint32_t* loffs,
                                 uint32_t* coffs, uint32_t* lit_total, uint32_t* cand_total, uint32_t* lit_snap,
                                 Rec* lout, uint32_t lcap, Rec* cout, uint32_t ccap, hipStream_t st) {
  hipLaunchKernelGGL(k_slice_offsets2, dim3(1), dim3(1024), 0, st, lit, cand, n, loffs, coffs, lit_total,
                     cand_total, lit_snap);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_slice_gather2, dim3(2 * (n + 64)), dim3(kTPB), 0, st, lit, cand, n, loffs, coffs, lout,
                     lcap, cout, ccap);
  return hipGetLastError();
}

// The literal slices straight to ordering pairs and bucket counts, and
// (cout set) the candidate slices to records, in one launch.
hipError_t launch_slice_compact_keys(const SliceOut& lit, const SliceOut& cand, uint32_t n, uint32_t* lit_total,
                                     uint32_t* cand_total, uint32_t* lit_snap, uint32_t cap, uint32_t sbits,
                                     uint32_t pbits, uint64_t* keys, uint32_t* vals, uint32_t* cnt, uint32_t shift,
                                     Rec* cout, uint32_t ccap, hipStream_t st, const DenseOut& dense) {
  hipLaunchKernelGGL(k_slice_gather_keys, dim3((cout ? 2 : 1) * (n + 64)), dim3(kTPB), 0, st, lit, n, cap, sbits,
                     pbits, keys, vals, cnt, shift, cand, cout, ccap, lit_total, cand_total, lit_snap, dense);
  return hipGetLastError();
}

// A batch's unit records from its runs ({item, first unit}, first units
// increasing from 0): unit u belongs to the last run whose first unit is
// <= u and is that item's tile (start & ~15) + (u - first) * unit.
__global__ void k_expand_units(const uint2* runs, uint32_t nruns, const uint2* items, uint32_t unit, uint4* units,
                               uint32_t nunits) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < nunits; u += gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = nruns;  // runs[lo].y <= u < runs[hi].y (hi = nruns: past the end)
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (runs[mid].y <= u) lo = mid;
      else hi = mid;
    }
    const uint2 r = runs[lo];
    const uint2 seg = items[r.x];
    units[u] = make_uint4((seg.x & ~15u) + (u - r.y) * unit, r.x, seg.x, seg.y);
  }
}

hipError_t launch_expand_units(const uint2* runs, uint32_t nruns, const uint2* items, uint32_t unit, uint4* units,
                               uint32_t nunits, hipStream_t st) {
  hipLaunchKernelGGL(k_expand_units, dim3(std::max(1u, std::min(4096u, (nunits + 255) / 256))), dim3(256), 0, st,
                     runs, nruns, items, unit, units, nunits);
  return hipGetLastError();
}

// Start of a device-sized call: zeroes the sticky words and the run
// counters, and resets the n {start, end} kernel stamp pairs (stamps may be
// null).
__global__ void k_call_init(uint32_t* sticky, uint32_t nsticky, uint32_t* counters, uint32_t ncounters,
                            unsigned long long* stamps, uint32_t n, CallItems ci, uint2* items) {
  const uint32_t t = threadIdx.x;
  if (t < nsticky) sticky[t] = 0;
  if (t < ncounters) counters[t] = 0;
  if (t < ci.n) items[t] = ci.v[t];
  if (stamps)
    for (uint32_t k = t; k < n; k += blockDim.x) {
      stamps[2 * k] = ~0ull;
      stamps[2 * k + 1] = 0;
    }
}

hipError_t launch_call_init(uint32_t* sticky, uint32_t nsticky, uint32_t* counters, uint32_t ncounters,
                            unsigned long long* stamps, uint32_t n, const CallItems& ci, uint2* items,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_call_init, dim3(1), dim3(256), 0, st, sticky, nsticky, counters, ncounters, stamps, n, ci,
                     items);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(kTPB), 0, st, p, n, v);
  return hipGetLastError();
}

// Stable LSD radix sort of (u64 key, u32 value) pairs over the low `bits` key bits.
hipError_t sort_pairs(void* tmp, size_t* tmp_bytes, const uint64_t* kin, uint64_t* kout,
                      const uint32_t* vin, uint32_t* vout, uint32_t n, uint32_t bits,
                      hipStream_t st) {
  return rocprim::radix_sort_pairs(tmp, *tmp_bytes, kin, kout, vin, vout, (size_t)n, 0u, bits, st);
}

// End of a device-sized run: record overflows and errors, accumulate
// statistics, advance the output offset by this run's match count, and zero
// the run's counters for the next run. Called by a whole block.
// fcnt: the filter's counter block (words 5, 6 and 8; zeroed here), or
// null when the filter counted in `counters`.
__device__ void run_close(uint32_t* counters, uint32_t* fcnt, const RunCaps& caps, uint32_t* sticky,
                          unsigned long long* total, uint32_t* publish, uint32_t pwords) {
  if (threadIdx.x == 0) {
    const uint32_t* f = fcnt ? fcnt : counters;
    uint32_t redo = 0;
    if (f[5] > caps.lit_ovf_cap || f[6] > caps.cand_ovf_cap) redo |= 2u;
    if (counters[kCounterLitSnapshot] > caps.lit_cap) redo |= 4u;
    if (counters[1] > caps.cand_cap) redo |= 8u;
    if (counters[2] > caps.raw_cap) redo |= 16u;
    if (counters[0] > caps.rec_cap) redo |= 32u;
    if (counters[9]) redo |= 64u;  // a sort bucket overflowed
    sticky[kStickyRedo] |= redo;
    if (counters[3] & kFlagTruncatedWalk) sticky[kStickyError] |= 1u;
    if (counters[0] > caps.max_matches) sticky[kStickyError] |= 2u;
    sticky[kStickyHits] += f[8];
    sticky[kStickyCands] += counters[1];
    sticky[kStickyRaw] += counters[2];
    *total += counters[0];
    for (uint32_t k = 0; k <= kCounterLitSnapshot; ++k) counters[k] = 0;
    if (fcnt) fcnt[5] = fcnt[6] = fcnt[8] = 0;
  }
  // The call's last run hands the state block (flags, total, kernel stamps)
  // to pinned host memory, so the host reads it without a copy.
  // publish[pwords] is then set to 1: the host may wait on it instead of
  // blocking on the stream.
  // It then leaves the state block as k_call_init sets it (flags and total
  // zero, stamp pairs {all ones, 0}), so a next call on the same stream
  // with the same segments may skip the init kernel.
  if (publish) {
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < pwords; k += blockDim.x) publish[k] = sticky[k];
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) {
      __hip_atomic_store(&publish[pwords], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    if (threadIdx.x < kStateResetWords) sticky[threadIdx.x] = 0;
    unsigned long long* stamps = reinterpret_cast<unsigned long long*>(sticky + 16);
    for (uint32_t k = threadIdx.x; k < (pwords - 16) / 2; k += blockDim.x) stamps[k] = (k & 1) ? 0ull : ~0ull;
  }
}

__global__ void k_run_close(uint32_t* counters, uint32_t* fcnt, RunCaps caps, uint32_t* sticky,
                            unsigned long long* total, uint32_t* publish, uint32_t pwords) {
  run_close(counters, fcnt, caps, sticky, total, publish, pwords);
}

__global__ void k_copy_u32(uint32_t* dst, const uint32_t* src) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *dst = *src;
}

hipError_t launch_run_close(uint32_t* counters, uint32_t* fcnt, const RunCaps& caps, uint32_t* sticky,
                            unsigned long long* total, uint32_t* publish, uint32_t pwords, hipStream_t st) {
  hipLaunchKernelGGL(k_run_close, dim3(1), dim3(64), 0, st, counters, fcnt, caps, sticky, total, publish, pwords);
  return hipGetLastError();
}

hipError_t launch_copy_u32(uint32_t* dst, const uint32_t* src, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_u32, dim3(1), dim3(64), 0, st, dst, src);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Bucket sort of (key, value) pairs whose keys spread over their range (match
// keys: position-major). Bucket = key >> shift, about two buckets per pair:
// count, exclusive scan (one block up to 2^16 buckets, else rocPRIM),
// scatter, then each bucket's pairs are ranked in LDS. A bucket over
// kBucketMax pairs is left unsorted and raises *flag; the caller then falls
// back to the radix sort.
// ---------------------------------------------------------------------------
constexpr uint32_t kBucketMax = 256;
constexpr int kSortWaves = 4;                     // buckets (one per wave) per block of k_bucket_rank

__global__ void k_bucket_count(const uint64_t* keys, uint32_t n, const uint32_t* nd, uint32_t shift,
                               uint32_t* cnt) {
  n = count_of(n, nd);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    atomicAdd(&cnt[keys[i] >> shift], 1u);
}

// Exclusive scan of nb <= kScanOneBlock counts into cursor in one 1024-lane
// block: tiles of 4096 counts, four per lane, all loaded up front (cnt and
// cursor hold kScanOneBlock entries, zero beyond nb); a wave scan per tile,
// the 16 x 16 wave totals scanned by wave 0 (tile-major).
// T tiles of 4096 counts (T * 4096 >= nb, T a power of two <= 16): the
// counts past nb are zero and not written.
template <int T>
__global__ __launch_bounds__(1024) void k_scan_counts(const uint32_t* cnt, uint32_t* cursor, uint32_t nb) {
  static_assert(T >= 1 && T <= kScanOneBlock / 4096 && 64 % T == 0, "tiles");
  __shared__ uint32_t s_tot[T * 16];
  const uint32_t tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const uint4* c4 = reinterpret_cast<const uint4*>(cnt);
  uint4 v[T];
  uint32_t own[T], inc[T];
#pragma unroll
  for (int t = 0; t < T; ++t) v[t] = c4[t * 1024 + tid];
#pragma unroll
  for (int t = 0; t < T; ++t) {
    own[t] = v[t].x + v[t].y + v[t].z + v[t].w;
    const uint32_t x = wave_incl_sum(own[t], lane);
    inc[t] = x;
    if (lane == 63) s_tot[t * 16 + w] = x;
  }
  __syncthreads();
  if (w == 0) {
    // The T * 16 wave totals (tile-major), T / 4 per lane: lane l holds
    // entries [l * T / 4, (l + 1) * T / 4) when T >= 4, else lanes below
    // T * 16 hold one each.
    constexpr int per = T >= 4 ? T / 4 : 1;
    uint32_t a[per];
    uint32_t sum = 0;
    const bool mine = T >= 4 || lane < (uint32_t)(T * 16);
#pragma unroll
    for (int k = 0; k < per; ++k) {
      a[k] = mine ? s_tot[lane * per + k] : 0u;
      sum += a[k];
    }
    const uint32_t x = wave_incl_sum(sum, lane);
    uint32_t ex = x - sum;
#pragma unroll
    for (int k = 0; k < per; ++k) {
      if (mine) s_tot[lane * per + k] = ex;
      ex += a[k];
    }
  }
  __syncthreads();
  uint4* o4 = reinterpret_cast<uint4*>(cursor);
#pragma unroll
  for (int t = 0; t < T; ++t) {
    if (t * 4096u + tid * 4u >= nb) continue;
    const uint32_t b = s_tot[t * 16 + w] + inc[t] - own[t];
    o4[t * 1024 + tid] = make_uint4(b, b + v[t].x, b + v[t].x + v[t].y, b + v[t].x + v[t].y + v[t].z);
  }
}

__global__ void k_bucket_scatter(const uint64_t* kin, const uint32_t* vin, uint32_t n, const uint32_t* nd,
                                 uint32_t shift, uint32_t* cursor, uint64_t* kout, uint32_t* vout) {
  n = count_of(n, nd);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint64_t k = kin[i];
    const uint32_t pos = atomicAdd(&cursor[k >> shift], 1u);
    kout[pos] = k;
    vout[pos] = vin[i];
  }
}

// Ranks each bucket's pairs (end[b] is its end after the scatter, cnt[b]
// its size): a pair's rank is the number of smaller keys, equal keys by
// index, so the order is stable. Two parts of one grid:
//  * blocks below `split`: a wave takes 64 consecutive buckets, one lane
//    each for those of 1..kLaneSort pairs, in the lane's own column of the
//    wave's LDS;
//  * the rest: wave k of W takes buckets k, k + W, ... (so runs of large
//    buckets spread over the waves) and ranks those over kLaneSort pairs
//    with the whole wave.
// Each part zeroes the counts it sorted (the counts start at zero for the
// next sort); a count one part reads is either its own or left alone by the
// other. Pairs are written back in place or, with PACK, as matches at the
// running output offset (put_match).
constexpr uint32_t kLaneSort = 8;
constexpr uint32_t kRankSlots = kLaneSort * 64 > kBucketMax ? kLaneSort * 64 : kBucketMax;

template <bool PACK>
__device__ __forceinline__ void rank_out(uint64_t* keys, uint32_t* vals, uint64_t slot, uint64_t k, uint32_t v,
                                         uint32_t pbits, uint32_t start_off, void* out, uint64_t out_cap,
                                         const DenseOut& dense) {
  if (PACK) {
    if (slot < out_cap) put_match(out, slot, k >> pbits, (uint32_t)(k & ((1ull << pbits) - 1)), v, start_off, dense);
  } else {
    keys[slot] = k;
    vals[slot] = v;
  }
}

template <bool PACK>
__global__ __launch_bounds__(64 * kSortWaves) void k_bucket_rank(uint64_t* keys, uint32_t* vals, const uint32_t* end,
                                                                  uint32_t* cnt, uint32_t nb, uint32_t split,
                                                                  uint32_t* flag, uint32_t pbits, uint32_t start_off,
                                                                  void* out, uint64_t out_cap,
                                                                  const unsigned long long* base_dev, DenseOut dense) {
  __shared__ uint64_t sk[kSortWaves][kRankSlots];
  __shared__ uint32_t sv[kSortWaves][kRankSlots];
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint64_t base = PACK ? *base_dev : 0ull;
  uint64_t* k_ = sk[w];
  uint32_t* v_ = sv[w];
  if (blockIdx.x < split) {
    for (uint32_t g = blockIdx.x * kSortWaves + w; g * 64u < nb; g += split * kSortWaves) {
      const uint32_t b = g * 64u + lane;
      const uint32_t c = b < nb ? cnt[b] : 0u;
      if (c == 0 || c > kLaneSort) continue;
      cnt[b] = 0;
      const uint32_t lo = end[b] - c;
      if (c == 1) {
        if (PACK) rank_out<PACK>(keys, vals, base + lo, keys[lo], vals[lo], pbits, start_off, out, out_cap, dense);
        continue;
      }
      // Column `lane`: entry t at t * 64 + lane.
      for (uint32_t t = 0; t < c; ++t) {
        k_[t * 64 + lane] = keys[lo + t];
        v_[t * 64 + lane] = vals[lo + t];
      }
      for (uint32_t t = 0; t < c; ++t) {
        const uint64_t k = k_[t * 64 + lane];
        uint32_t r = 0;
        for (uint32_t u = 0; u < c; ++u) {
          const uint64_t ku = k_[u * 64 + lane];
          r += (ku < k) || (ku == k && u < t);
        }
        rank_out<PACK>(keys, vals, base + lo + r, k, v_[t * 64 + lane], pbits, start_off, out, out_cap, dense);
      }
    }
    return;
  }
  const uint32_t W = (gridDim.x - split) * kSortWaves;
  const uint32_t k0 = (blockIdx.x - split) * kSortWaves + w;
  // Wave-uniform loop: 64 of the wave's buckets per round, one per lane.
  for (uint32_t m = 0; (uint64_t)k0 + (uint64_t)W * 64u * m < nb; ++m) {
    const uint64_t b = (uint64_t)k0 + (uint64_t)W * (64u * m + lane);
    const uint32_t c = b < nb ? cnt[b] : 0u;
    uint64_t big = __ballot(c > kLaneSort);
    while (big) {
      const uint32_t l = __builtin_ctzll(big);
      big &= big - 1;
      const uint32_t cb = __builtin_amdgcn_readlane(c, l);
      const uint32_t bb = __builtin_amdgcn_readlane((uint32_t)b, l);
      if (lane == 0) cnt[bb] = 0;
      if (cb > kBucketMax) {
        if (lane == 0) atomicOr(flag, 1u);
        continue;
      }
      const uint32_t lob = end[bb] - cb;
      for (uint32_t i = lane; i < cb; i += 64) {
        k_[i] = keys[lob + i];
        v_[i] = vals[lob + i];
      }
      wave_lds_sync();
      for (uint32_t i = lane; i < cb; i += 64) {
        const uint64_t k = k_[i];
        uint32_t r = 0;
        for (uint32_t j = 0; j < cb; ++j) {
          const uint64_t kj = k_[j];
          r += (kj < k) || (kj == k && j < i);
        }
        rank_out<PACK>(keys, vals, base + lob + r, k, v_[i], pbits, start_off, out, out_cap, dense);
      }
      wave_lds_sync();  // the next bucket reuses the slots
    }
  }
}

// cursor = exclusive scan of cnt (nb counts).
static hipError_t scan_counts(void* tmp, size_t tmp_bytes, uint32_t* cnt, uint32_t* cursor, uint32_t nb,
                              hipStream_t st) {
  if (nb <= kScanOneBlock) {
    const uint32_t tiles = (nb + 4095) / 4096;
    if (tiles <= 1) hipLaunchKernelGGL(k_scan_counts<1>, dim3(1), dim3(1024), 0, st, cnt, cursor, nb);
    else if (tiles <= 2) hipLaunchKernelGGL(k_scan_counts<2>, dim3(1), dim3(1024), 0, st, cnt, cursor, nb);
    else if (tiles <= 4) hipLaunchKernelGGL(k_scan_counts<4>, dim3(1), dim3(1024), 0, st, cnt, cursor, nb);
    else if (tiles <= 8) hipLaunchKernelGGL(k_scan_counts<8>, dim3(1), dim3(1024), 0, st, cnt, cursor, nb);
    else hipLaunchKernelGGL(k_scan_counts<16>, dim3(1), dim3(1024), 0, st, cnt, cursor, nb);
    return hipGetLastError();
  }
  return rocprim::exclusive_scan(tmp, tmp_bytes, cnt, cursor, 0u, (size_t)nb, rocprim::plus<uint32_t>(), st);
}

// Blocks of k_bucket_rank: `split` of lane-sorting blocks, then the
// wave-sorting ones (one wave per 16 buckets, at most 1024 blocks).
static void rank_grid(uint32_t nb, uint32_t* split, uint32_t* grid) {
  const uint32_t groups = (nb + 63) / 64;
  *split = std::max(1u, std::min(2048u, (groups + kSortWaves - 1) / kSortWaves));
  *grid = *split + std::max(1u, std::min(1024u, (nb / 16 + kSortWaves - 1) / kSortWaves));
}
// ---------------------------------------------------------------------------
// Ordering of a literal-only set's matches in three launches. A match's key is
// start << pbits | pattern id (a literal-only set's pattern ids are its
// literal indices, so the end is start + the literal's length). The
// position range is cut into nb buckets of kLitBucketCap slots.
//  1. k_gather_buckets: every slice record to a slot of its bucket (a
//     bucket that fills up sets the overflow flag: the call is redone on
//     the checked path).
//  2. k_scan_buckets: the bucket counts to output offsets, and the total.
//  3. k_rank_pack: one wave per bucket ranks its keys in LDS and writes the
//     matches in order; it leaves the counts zero.
// ---------------------------------------------------------------------------
constexpr int kGatherTPB = 1024;

__global__ __launch_bounds__(kGatherTPB) void k_gather_buckets(SliceOut lit, uint32_t n, uint32_t pbits,
                                                               uint32_t shift, uint64_t* bkeys, uint32_t* bcnt,
                                                               uint32_t* counters, DenseOut dense) {
  // Called by whole waves (`live` false for lanes without a record). Lanes
  // of a wave holding the same bucket share one atomic: consecutive slice
  // records mostly fall in one or two buckets.
  auto put = [&](bool live, const Rec& r) {
    const uint64_t k = ((dense.bases && live ? dense.bases[r.w] + r.y : (uint64_t)r.y) << pbits) | r.x;
    const uint32_t b = (uint32_t)(k >> shift);
    uint64_t todo = __ballot(live);
    while (todo) {
      const int leader = __ffsll((unsigned long long)todo) - 1;
      const uint32_t lb = __builtin_amdgcn_readlane(b, leader);
      const uint64_t same = __ballot(live && b == lb) & todo;
      uint32_t base = 0;
      if ((int)lane_id() == leader) base = atomicAdd(&bcnt[lb], (uint32_t)__popcll(same));
      base = __builtin_amdgcn_readlane(base, leader);
      if ((same >> lane_id()) & 1ull) {
        const uint32_t slot = base + (uint32_t)__popcll(same & ((1ull << lane_id()) - 1ull));
        if (slot < kLitBucketCap)
          bkeys[(size_t)b * kLitBucketCap + slot] = k;
        else
          atomicOr(&counters[9], 1u);  // a bucket is full: redo on the checked path
      }
      todo &= ~same;
    }
  };
  if (blockIdx.x < n) {
    const uint32_t m = min(lit.counts[blockIdx.x], lit.cap);
    for (uint32_t k0 = 0; k0 < m; k0 += kGatherTPB) {
      const uint32_t k = k0 + threadIdx.x;
      const bool live = k < m;
      put(live, live ? lit.buf[(size_t)blockIdx.x * lit.cap + k] : make_uint4(0, 0, 0, 0));
    }
  } else {
    const uint32_t no = min(*lit.ovf_count, lit.ovf_cap);
    for (uint32_t k0 = (blockIdx.x - n) * kGatherTPB; k0 < no; k0 += 64 * kGatherTPB) {
      const uint32_t k = k0 + threadIdx.x;
      const bool live = k < no;
      put(live, live ? lit.ovf[k] : make_uint4(0, 0, 0, 0));
    }
  }
}

// Exclusive scan of the kLitBuckets bucket counts (each clamped to the
// bucket capacity) into output offsets, and the match total. One block.
__global__ __launch_bounds__(kGatherTPB) void k_scan_buckets(const uint32_t* bcnt, uint32_t* bofs,
                                                             uint32_t* counters) {
  __shared__ uint32_t s_wt[kGatherTPB / 64];
  constexpr uint32_t per = kLitBuckets / kGatherTPB;
  const uint32_t b0 = threadIdx.x * per;
  uint32_t v[per];
#pragma unroll
  for (uint32_t k = 0; k < per; k += 4) {
    const uint4 x = *reinterpret_cast<const uint4*>(bcnt + b0 + k);
    v[k] = min(x.x, kLitBucketCap);
    v[k + 1] = min(x.y, kLitBucketCap);
    v[k + 2] = min(x.z, kLitBucketCap);
    v[k + 3] = min(x.w, kLitBucketCap);
  }
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < per; ++k) sum += v[k];
  // Block prefix: a DPP wave prefix, the wave totals through LDS (one
  // barrier).
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t incl = wave_incl_sum(sum, lane);
  if (lane == 63) s_wt[w] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (uint32_t k = 0; k < kGatherTPB / 64; ++k) {
    const uint32_t t = s_wt[k];
    before += k < w ? t : 0u;
    all += t;
  }
  uint32_t run = before + incl - sum;
#pragma unroll
  for (uint32_t k = 0; k < per; k += 4) {
    uint4 o;
    o.x = run; run += v[k];
    o.y = run; run += v[k + 1];
    o.z = run; run += v[k + 2];
    o.w = run; run += v[k + 3];
    *reinterpret_cast<uint4*>(bofs + b0 + k) = o;
  }
  if (threadIdx.x == kGatherTPB - 1) {
    counters[0] = all;
    counters[kCounterLitSnapshot] = all;
  }
}

constexpr int kRankWaves = 4;  // buckets (one per wave) per block of k_rank_pack

__global__ __launch_bounds__(64 * kRankWaves) void k_rank_pack(const uint64_t* bkeys, uint32_t* bcnt,
                                                               const uint32_t* bofs, uint32_t nb, uint32_t pbits,
                                                               const uint32_t* flen, uint32_t start_off,
                                                               void* out, uint64_t out_cap,
                                                               const unsigned long long* base_dev, DenseOut dense) {
  __shared__ uint64_t s_keys[kRankWaves][kLitBucketCap];
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t b = blockIdx.x * kRankWaves + w;
  if (b >= nb) return;
  const uint32_t c = min(bcnt[b], kLitBucketCap);
  if (c == 0) return;
  const uint64_t* src = bkeys + (size_t)b * kLitBucketCap;
  for (uint32_t i = lane; i < c; i += 64) s_keys[w][i] = src[i];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const uint64_t base = *base_dev + bofs[b];
  const uint64_t pmask = (1ull << pbits) - 1;
  // Keys are distinct (one match per pattern and start), so a key's rank is
  // the number of smaller keys.
  for (uint32_t i = lane; i < c; i += 64) {
    const uint64_t k = s_keys[w][i];
    uint32_t r = 0;
    for (uint32_t j = 0; j < c; ++j) r += s_keys[w][j] < k;
    if (base + r < out_cap) {
      const uint32_t pid = (uint32_t)(k & pmask);
      put_match(out, base + r, k >> pbits, pid, flen[pid], start_off, dense, true);
    }
  }
  if (lane == 0) bcnt[b] = 0;
}

hipError_t launch_lit_order(const SliceOut& lit, uint32_t n, uint32_t pbits, uint32_t shift, uint32_t nb,
                            uint64_t* bkeys, uint32_t* bcnt, uint32_t* bofs, uint32_t* counters,
                            const uint32_t* flen, uint32_t start_off, void* out, uint64_t out_cap,
                            const unsigned long long* base_dev, hipStream_t st, const DenseOut& dense) {
  hipLaunchKernelGGL(k_gather_buckets, dim3(n + 64), dim3(kGatherTPB), 0, st, lit, n, pbits, shift, bkeys, bcnt,
                     counters, dense);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_buckets, dim3(1), dim3(kGatherTPB), 0, st, bcnt, bofs, counters);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_rank_pack, dim3((nb + kRankWaves - 1) / kRankWaves), dim3(64 * kRankWaves), 0, st, bkeys,
                     bcnt, bofs, nb, pbits, flen, start_off, out, out_cap, base_dev, dense);
  return hipGetLastError();
}

// The same ordering for a run with at most kSmallOrder matches, and the
// run's close, in one block: the slices' records (and the overflow area's)
// are counted into kSmallBuckets position buckets in LDS, scattered by
// bucket, ranked within their bucket and packed in order. More matches, or
// a bucket over kLitBucketCap, set the bucket-overflow flag: the call is
// redone on the checked path, and the host returns to the three-launch
// ordering. One launch instead of four: on a 16 MiB input with a few
// hundred matches the kernels after the filter took 16 us, most of it the
// fixed cost of each launch.
constexpr uint32_t kSmallOrderTPB = 1024;
constexpr uint32_t kSmallBuckets = 1024;
constexpr uint32_t kSmallPer = kSmallOrder / kSmallOrderTPB;  // records per lane

// Exclusive prefix sum over the block (one value per lane, 1024 lanes);
// *total gets the sum. s_w: 16 words of scratch.
__device__ __forceinline__ uint32_t block_excl_sum(uint32_t v, uint32_t* s_w, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t x = wave_incl_sum(v, lane);
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (uint32_t k = 0; k < kSmallOrderTPB / 64; ++k) {
    const uint32_t t = s_w[k];
    before += k < w ? t : 0u;
    all += t;
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

__global__ __launch_bounds__(kSmallOrderTPB) void k_lit_order_small(
    SliceOut lit, uint32_t n, uint32_t pbits, uint32_t bshift, uint32_t start_off, void* out, uint64_t out_cap,
    uint32_t* counters, uint32_t* fcnt, RunCaps caps, uint32_t* sticky, unsigned long long* total,
    uint32_t* publish, uint32_t pwords) {
  static_assert(kSmallOrderSlices + 1 <= kSmallOrderTPB && kSmallBuckets == kSmallOrderTPB, "one per lane");
  __shared__ uint64_t s_keys[kSmallOrder];
  __shared__ uint32_t s_end[kSmallOrder];
  __shared__ uint32_t s_off[kSmallOrderTPB + 1];
  __shared__ uint32_t s_bcnt[kSmallBuckets];
  __shared__ uint32_t s_bofs[kSmallBuckets];
  __shared__ uint32_t s_w[kSmallOrderTPB / 64];
  __shared__ uint32_t s_state[kStateResetWords];
  __shared__ unsigned long long s_base;
  const uint32_t tid = threadIdx.x;
  // The close's inputs, read first so that their latency overlaps the
  // ordering: the filter's counters, the state block (and, for the host
  // copy, its kernel stamps) and the running total.
  uint32_t cnt_in[kCounterLitSnapshot + 1] = {};
  uint32_t st_in[5] = {};
  if (tid == 0) {
#pragma unroll
    for (uint32_t k = 0; k <= kCounterLitSnapshot; ++k) cnt_in[k] = counters[k];
    if (fcnt) {  // the filter counted its hits and overflows in its own block
      cnt_in[5] = fcnt[5];
      cnt_in[6] = fcnt[6];
      cnt_in[8] = fcnt[8];
    }
#pragma unroll
    for (uint32_t k = 0; k < 5; ++k) st_in[k] = sticky[k];
    s_base = *total;
  }
  const uint32_t pub_in = publish && tid < pwords ? sticky[tid] : 0u;
  // Slice record counts to offsets (slice `tid`; lane n: the overflow area).
  uint32_t c = 0;
  if (tid < n) c = min(lit.counts[tid], lit.cap);
  else if (tid == n) c = min(*lit.ovf_count, lit.ovf_cap);
  uint32_t T = 0;
  s_off[tid] = block_excl_sum(c, s_w, &T);
  s_bcnt[tid] = 0;
  if (tid == 0) s_off[kSmallOrderTPB] = T;
  __syncthreads();
  bool ok = T <= kSmallOrder;
  uint64_t key[kSmallPer];
  uint32_t slot[kSmallPer], endv[kSmallPer];
  if (ok) {
#pragma unroll
    for (uint32_t q = 0; q < kSmallPer; ++q) {
      const uint32_t i = tid + q * kSmallOrderTPB;
      if (i < T) {
        uint32_t lo = 0, hi = n + 1;  // s_off[lo] <= i < s_off[hi]
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) / 2;
          if (s_off[mid] <= i) lo = mid;
          else hi = mid;
        }
        const Rec r = lo < n ? lit.buf[(size_t)lo * lit.cap + (i - s_off[lo])] : lit.ovf[i - s_off[n]];
        key[q] = ((uint64_t)r.y << pbits) | r.x;
        endv[q] = r.z;  // a literal match record carries its end
        slot[q] = atomicAdd(&s_bcnt[min((uint32_t)(r.y >> bshift), kSmallBuckets - 1)], 1u);
      }
    }
  }
  __syncthreads();
  const uint32_t bc = s_bcnt[tid];
  uint32_t unused = 0;
  s_bofs[tid] = block_excl_sum(bc, s_w, &unused);
  // A crowded bucket would cost a quadratic rank: the checked path instead.
  if (bc > kLitBucketCap) ok = false;
  const bool fail = __syncthreads_or(!ok);
  if (!fail) {
#pragma unroll
    for (uint32_t q = 0; q < kSmallPer; ++q) {
      const uint32_t i = tid + q * kSmallOrderTPB;
      if (i < T) {
        const uint32_t b = min((uint32_t)((key[q] >> pbits) >> bshift), kSmallBuckets - 1);
        s_keys[s_bofs[b] + slot[q]] = key[q];
        s_end[s_bofs[b] + slot[q]] = endv[q];
      }
    }
    __syncthreads();
    // Keys are distinct (one match per pattern and start): a key's place is
    // its bucket's offset plus the number of smaller keys in its bucket.
    const uint64_t pmask = (1ull << pbits) - 1, base = s_base;
#pragma unroll
    for (uint32_t q = 0; q < kSmallPer; ++q) {
      const uint32_t i = tid + q * kSmallOrderTPB;
      if (i < T) {
        const uint64_t k = s_keys[i];
        const uint32_t b = min((uint32_t)((k >> pbits) >> bshift), kSmallBuckets - 1);
        const uint32_t b0 = s_bofs[b], b1 = b0 + s_bcnt[b];
        uint32_t r = 0;
        for (uint32_t j = b0; j < b1; ++j) r += s_keys[j] < k;
        if (base + b0 + r < out_cap) {
          MatchOut mo;
          mo.pid = (uint32_t)(k & pmask);
          mo.start = (uint32_t)(k >> pbits) + start_off;
          mo.end = s_end[i] + start_off;
          reinterpret_cast<MatchOut*>(out)[base + b0 + r] = mo;
        }
      }
    }
  }
  // The close (what run_close does, from the values read at the start):
  // this run's count is T, or it is redone when it did not fit.
  if (tid == 0) {
    const uint32_t c0 = fail ? 0u : T;
    uint32_t redo = 0;
    if (cnt_in[5] > caps.lit_ovf_cap || cnt_in[6] > caps.cand_ovf_cap) redo |= 2u;
    if (c0 > caps.lit_cap) redo |= 4u;
    if (cnt_in[1] > caps.cand_cap) redo |= 8u;
    if (cnt_in[2] > caps.raw_cap) redo |= 16u;
    if (c0 > caps.rec_cap) redo |= 32u;
    if (fail || cnt_in[9]) redo |= 64u;
    uint32_t e = st_in[kStickyError];
    if (cnt_in[3] & kFlagTruncatedWalk) e |= 1u;
    if (c0 > caps.max_matches) e |= 2u;
    const unsigned long long tot = s_base + c0;
    s_state[kStickyRedo] = st_in[kStickyRedo] | redo;
    s_state[kStickyError] = e;
    s_state[kStickyHits] = st_in[kStickyHits] + cnt_in[8];
    s_state[kStickyCands] = st_in[kStickyCands] + cnt_in[1];
    s_state[kStickyRaw] = st_in[kStickyRaw] + cnt_in[2];
    s_state[5] = s_state[6] = s_state[7] = 0;
    s_state[8] = (uint32_t)tot;
    s_state[9] = (uint32_t)(tot >> 32);
#pragma unroll
    for (uint32_t k = 0; k < 5; ++k) sticky[k] = s_state[k];
    *total = tot;
#pragma unroll
    for (uint32_t k = 0; k <= kCounterLitSnapshot; ++k) counters[k] = 0;
    if (fcnt) fcnt[5] = fcnt[6] = fcnt[8] = 0;
  }
  if (publish) {
    // The state block to pinned host memory (words 0-9 as just closed, the
    // stamps as read), then the completion flag; then the block is left as
    // the call's init kernel sets it.
    __syncthreads();
    if (tid < pwords) publish[tid] = tid < kStateResetWords ? s_state[tid] : pub_in;
    for (uint32_t k = tid + kSmallOrderTPB; k < pwords; k += kSmallOrderTPB) publish[k] = sticky[k];
    __threadfence_system();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&publish[pwords], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    if (tid < kStateResetWords) sticky[tid] = 0;
    unsigned long long* stamps = reinterpret_cast<unsigned long long*>(sticky + 16);
    for (uint32_t k = tid; k < (pwords - 16) / 2; k += kSmallOrderTPB) stamps[k] = (k & 1) ? 0ull : ~0ull;
  }
}

// sbits: bits of the largest start (the run's segment length).
hipError_t launch_lit_order_small(const SliceOut& lit, uint32_t n, uint32_t pbits, uint32_t sbits,
                                  const uint32_t* flen, uint32_t start_off, void* out, uint64_t out_cap,
                                  uint32_t* counters, uint32_t* fcnt, const RunCaps& caps, uint32_t* sticky,
                                  unsigned long long* total, uint32_t* publish, uint32_t pwords, hipStream_t st) {
  if (n > kSmallOrderSlices) return hipErrorInvalidValue;
  const uint32_t bshift = sbits > 10 ? sbits - 10 : 0;  // kSmallBuckets = 2^10
  (void)flen;  // a literal match record carries its end
  hipLaunchKernelGGL(k_lit_order_small, dim3(1), dim3(kSmallOrderTPB), 0, st, lit, n, pbits, bshift, start_off,
                     out, out_cap, counters, fcnt, caps, sticky, total, publish, pwords);
  return hipGetLastError();
}

// Pairs already counted into cnt (k_slice_gather_keys): scan -> scatter ->
// per-bucket rank fused with packing into `out` at *base_dev.
hipError_t bucket_order_pack(void* tmp, size_t tmp_bytes, const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
                             uint32_t* vout, uint32_t n, const uint32_t* nd, uint32_t shift, uint32_t nb,
                             uint32_t* cnt, uint32_t* cursor, uint32_t* flag, uint32_t pbits, uint32_t start_off,
                             void* out, uint64_t out_cap, const unsigned long long* base_dev, hipStream_t st,
                             const DenseOut& dense) {
  hipError_t e = scan_counts(tmp, tmp_bytes, cnt, cursor, nb, st);
  if (e != hipSuccess) return e;
  const uint32_t g = std::max(1u, std::min(nd ? kDeviceGrid : 4096u, (n + 255) / 256));
  hipLaunchKernelGGL(k_bucket_scatter, dim3(g), dim3(256), 0, st, kin, vin, n, nd, shift, cursor, kout, vout);
  uint32_t split, grid;
  rank_grid(nb, &split, &grid);
  hipLaunchKernelGGL(k_bucket_rank<true>, dim3(grid), dim3(64 * kSortWaves), 0, st, kout, vout, cursor, cnt, nb,
                     split, flag, pbits, start_off, out, out_cap, base_dev, dense);
  return hipGetLastError();
}

// Sizes the scan's temporary storage (tmp = nullptr) or runs
// count -> scan -> scatter -> per-bucket sort (with `counted`, the pairs'
// producer has counted them already). cnt and cursor hold nb words; cnt must
// be zero on entry (apart from a producer's counts) and is zero again on
// exit. With nd set, n is a capacity and the count is *nd.
hipError_t bucket_sort_pairs(void* tmp, size_t* tmp_bytes, const uint64_t* kin, uint64_t* kout,
                             const uint32_t* vin, uint32_t* vout, uint32_t n, const uint32_t* nd, uint32_t shift,
                             uint32_t nb, uint32_t* cnt, uint32_t* cursor, uint32_t* flag, hipStream_t st,
                             bool counted) {
  if (!tmp)
    return rocprim::exclusive_scan(nullptr, *tmp_bytes, cnt, cursor, 0u, (size_t)nb, rocprim::plus<uint32_t>(),
                                   st);
  const uint32_t g = std::max(1u, std::min(nd ? kDeviceGrid : 4096u, (n + 255) / 256));
  if (!counted) hipLaunchKernelGGL(k_bucket_count, dim3(g), dim3(256), 0, st, kin, n, nd, shift, cnt);
  hipError_t e = scan_counts(tmp, *tmp_bytes, cnt, cursor, nb, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bucket_scatter, dim3(g), dim3(256), 0, st, kin, vin, n, nd, shift, cursor, kout, vout);
  uint32_t split, grid;
  rank_grid(nb, &split, &grid);
  hipLaunchKernelGGL(k_bucket_rank<false>, dim3(grid), dim3(64 * kSortWaves), 0, st, kout, vout, cursor, cnt, nb,
                     split, flag, 0u, 0u, nullptr, 0ull, nullptr, DenseOut{});
  return hipGetLastError();
}

// Explicit instantiations used by the host runtime.
template __global__ void k_ac_scan<true, 64>(ScanCommon, AcParams);
template __global__ void k_ac_scan<false, 64>(ScanCommon, AcParams);

}  // namespace ws
