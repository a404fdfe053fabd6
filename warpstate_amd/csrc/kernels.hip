// HIP kernels of the scan path, written for gfx950 (CDNA4, wave64).
//
//  k_hash_scan   rolling-hash prefilter over 16 positions per lane, bloom
//                filter in LDS, bucket chains in L2, exact verify, wave
//                prefix-sum compaction into an LDS staging buffer.
//  k_ac_scan     dense factor-DFA (Aho-Corasick over byte classes) traversal,
//                one chunk per lane, table in LDS or L2.
//  k_rx_walk     anchored leftmost-first DFA walk from each regex candidate.
//  k_chain_*     per-group greedy non-overlap chains (find_iter semantics).
//  k_*_keys/pack ordering keys and output packing.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include "kernels.hpp"
#include "regex.hpp"

namespace ws {

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// (a mod 2^24) * (b mod 2^24) + c, one v_mad_u32_u24.
__device__ __forceinline__ uint32_t mad24(uint32_t a, uint32_t b, uint32_t c) {
  return __umul24(a, b) + c;
}

// Same, pinned to a single v_mad_u32_u24 with a uniform multiplier (the
// compiler otherwise splits chained multiply-adds into mul + mul + add3).
__device__ __forceinline__ uint32_t mad24_s(uint32_t a, uint32_t b_uniform, uint32_t c) {
  uint32_t r;
  asm volatile("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b_uniform), "v"(c));
  return r;
}

__device__ __forceinline__ uint8_t fold_byte(uint8_t b) {
  return (b >= 'A' && b <= 'Z') ? (uint8_t)(b + 32) : b;
}

// SWAR ASCII lower-casing of four bytes.
__device__ __forceinline__ uint32_t fold_word(uint32_t x) {
  uint32_t hi = x & 0x80808080u;
  uint32_t v = x & 0x7F7F7F7Fu;
  uint32_t ge_a = v + 0x3F3F3F3Fu;   // byte >= 'A' (0x41) sets bit 7
  uint32_t gt_z = v + 0x25252525u;   // byte >  'Z' (0x5A) sets bit 7
  uint32_t upper = ge_a & ~gt_z & ~hi & 0x80808080u;
  return x | (upper >> 2);
}

// Exclusive wave prefix of `n` plus one LDS reservation for the wave total.
// Must be called with the whole wave active.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t n, uint32_t* counter) {
  const uint32_t lane = lane_id();
  uint32_t x = n;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= (uint32_t)d) x += y;
  }
  uint32_t total = __shfl(x, 63, 64);
  uint32_t base = 0;
  if (lane == 63 && total) base = atomicAdd(counter, total);
  base = __shfl(base, 63, 64);
  return base + x - n;
}

__device__ __forceinline__ bool verify_factor(const ScanCommon& c, uint64_t q, uint32_t hi,
                                              uint32_t f) {
  const uint32_t L = c.flen[f];
  if (q + L > hi) return false;
  const uint8_t* pb = c.fbytes + c.foff[f];
  const uint8_t* in = c.base + q;
  for (uint32_t t = 0; t < L; ++t) {
    uint8_t b = in[t];
    if (c.ci) b = fold_byte(b);
    if (b != pb[t]) return false;
  }
  return true;
}

__device__ __forceinline__ void stage_rec(Rec r, uint32_t slot, Rec* lds, uint32_t* gcount,
                                          Rec* gout, uint32_t gcap) {
  if (slot < kRecCap) {
    lds[slot] = r;
  } else {
    uint32_t g = atomicAdd(gcount, 1u);
    if (g < gcap) gout[g] = r;
  }
}

// Copies the block's staged records to global memory with one atomic per list.
__device__ void flush_block(uint32_t* s_n, Rec* s_lit, Rec* s_cand, uint32_t* s_base,
                            const ScanCommon& c) {
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t nl = min(s_n[0], kRecCap), nc = min(s_n[1], kRecCap);
    s_base[0] = nl ? atomicAdd(&c.counters[0], nl) : 0;
    s_base[1] = nc ? atomicAdd(&c.counters[1], nc) : 0;
    s_base[2] = nl;
    s_base[3] = nc;
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < s_base[2]; k += blockDim.x) {
    uint32_t g = s_base[0] + k;
    if (g < c.lit_cap) c.lit_out[g] = s_lit[k];
  }
  for (uint32_t k = threadIdx.x; k < s_base[3]; k += blockDim.x) {
    uint32_t g = s_base[1] + k;
    if (g < c.cand_cap) c.cand_out[g] = s_cand[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) { s_n[0] = 0; s_n[1] = 0; }
  __syncthreads();
}

__device__ __forceinline__ uint32_t byte_of(const uint32_t* d, int j) {
  return (d[j >> 2] >> ((j & 3) * 8)) & 0xFFu;
}

// ---------------------------------------------------------------------------
// Rolling-hash prefilter. Each lane owns 16 consecutive positions; the K-byte
// window hash is H(q) = sum b[q+t] B^(K-1-t) mod 2^24, rolled with two 24-bit
// multiply-adds per position. A set bit in the LDS bloom filter sends the
// position to the bucket chain (L2) and exact byte verification.
// ---------------------------------------------------------------------------
template <int K, bool CI>
__global__ __launch_bounds__(kTPB) void k_hash_scan(ScanCommon c, HashParams h) {
  extern __shared__ uint32_t smem[];
  uint32_t* bloom = smem;
  Rec* s_lit = reinterpret_cast<Rec*>(smem + h.nwords);
  Rec* s_cand = s_lit + kRecCap;
  __shared__ uint32_t s_n[2];
  __shared__ uint32_t s_base[4];

  for (uint32_t k = threadIdx.x; k < h.nwords; k += kTPB) bloom[k] = h.bloom[k];
  if (threadIdx.x == 0) { s_n[0] = 0; s_n[1] = 0; }
  __syncthreads();

  for (uint32_t u = blockIdx.x; u < c.nunits; u += gridDim.x) {
    uint64_t tile_base;
    uint32_t item;
    if (c.units) {
      uint2 uu = c.units[u];
      tile_base = uu.x;
      item = uu.y;
    } else {
      tile_base = (uint64_t)u * kHashUnit;
      item = 0;
    }
    const uint2 seg = c.items[item];
    const uint64_t lo = seg.x, hi = seg.y;
    const uint64_t p0 = tile_base + 16ull * threadIdx.x;
    uint32_t mask = 0;
    if (p0 < hi && p0 + 16 > lo && hi >= (uint64_t)K) {
      uint32_t d[8];
      uint64_t lend = (hi + 15) & ~15ull;
      if (lend > c.load_end) lend = c.load_end;
      if (p0 + 32 <= lend) {
        uint4 a = *reinterpret_cast<const uint4*>(c.base + p0);
        uint4 b = *reinterpret_cast<const uint4*>(c.base + p0 + 16);
        d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w;
        d[4] = b.x; d[5] = b.y; d[6] = b.z; d[7] = b.w;
      } else {
        uint4 a = *reinterpret_cast<const uint4*>(c.base + p0);
        d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w;
        d[4] = d[5] = d[6] = d[7] = 0;
      }
      if (CI) {
#pragma unroll
        for (int j = 0; j < 8; ++j) d[j] = fold_word(d[j]);
      }
      uint32_t H = 0;
#pragma unroll
      for (int t = 0; t < K; ++t) H = mad24(H, h.B, byte_of(d, t));
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        if (i > 0) H = mad24_s(H, h.B, mad24_s(byte_of(d, i - 1), h.NBK, byte_of(d, i + K - 1)));
        const uint32_t w = bloom[__builtin_amdgcn_ubfe(H, h.wshift, h.wbits)];
        // v_bfe_u32 reads its offset from H[4:0].
        mask |= __builtin_amdgcn_ubfe(w, H, 1) << i;
      }
      // Positions q are valid when lo <= q, q + K <= hi and q - lo < emit_end.
      uint64_t qmin = p0 > lo ? p0 : lo;
      uint64_t qmax = hi - K + 1;
      uint64_t em = lo + c.emit_end;
      if (em < qmax) qmax = em;
      if (qmax > p0 + 16) qmax = p0 + 16;
      uint32_t valid = 0;
      if (qmax > qmin) {
        uint32_t a0 = (uint32_t)(qmin - p0), a1 = (uint32_t)(qmax - p0);
        valid = ((a1 >= 32 ? 0xFFFFFFFFu : ((1u << a1) - 1u))) & ~((1u << a0) - 1u);
      }
      mask &= valid;
    }
    // Count pass over the (rare) filter hits, then one reservation per wave.
    uint32_t nl = 0, nc = 0;
    for (uint32_t m = mask; m; m &= m - 1) {
      const uint64_t q = p0 + __builtin_ctz(m);
      uint32_t H = 0;
      for (int t = 0; t < K; ++t) {
        uint8_t b = c.base[q + t];
        if (CI) b = fold_byte(b);
        H = (H * h.B + b) & 0xFFFFFFu;
      }
      const uint32_t bk = (H * 0x9E3779B1u) >> (32 - h.nbits);
      for (uint32_t e = h.bstart[bk]; e < h.bstart[bk + 1]; ++e) {
        const uint32_t f = h.bent[e];
        if (h.fhash[f] != H || !verify_factor(c, q, (uint32_t)hi, f)) continue;
        if (c.fown[f] >> 31) ++nc; else ++nl;
      }
    }
    if (__any((nl | nc) != 0u)) {
      uint32_t lslot = wave_reserve(nl, &s_n[0]);
      uint32_t cslot = wave_reserve(nc, &s_n[1]);
      for (uint32_t m = (nl | nc) ? mask : 0u; m; m &= m - 1) {
        const uint64_t q = p0 + __builtin_ctz(m);
        uint32_t H = 0;
        for (int t = 0; t < K; ++t) {
          uint8_t b = c.base[q + t];
          if (CI) b = fold_byte(b);
          H = (H * h.B + b) & 0xFFFFFFu;
        }
        const uint32_t bk = (H * 0x9E3779B1u) >> (32 - h.nbits);
        for (uint32_t e = h.bstart[bk]; e < h.bstart[bk + 1]; ++e) {
          const uint32_t f = h.bent[e];
          if (h.fhash[f] != H || !verify_factor(c, q, (uint32_t)hi, f)) continue;
          const uint32_t own = c.fown[f];
          const uint32_t start = (uint32_t)(q - lo);
          if (own >> 31) {
            stage_rec(make_uint4(own & 0x7FFFFFFFu, start, 0u, item), cslot++, s_cand,
                      &c.counters[1], c.cand_out, c.cand_cap);
          } else {
            stage_rec(make_uint4(c.lit_pid[own], start, start + c.flen[f], item), lslot++, s_lit,
                      &c.counters[0], c.lit_out, c.lit_cap);
          }
        }
      }
    }
    __syncthreads();
    if (s_n[0] > kRecCap / 2 || s_n[1] > kRecCap / 2) flush_block(s_n, s_lit, s_cand, s_base, c);
  }
  flush_block(s_n, s_lit, s_cand, s_base, c);
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

  for (uint32_t k = threadIdx.x; k < 256; k += kTPB) cmap[k] = a.cmap[k];
  if (LDS_TABLE) {
    const uint32_t n = a.states * a.classes;
    for (uint32_t k = threadIdx.x; k < n; k += kTPB) table[k] = a.table16[k];
  }
  if (threadIdx.x == 0) { s_n[0] = 0; s_n[1] = 0; }
  __syncthreads();

  for (uint32_t u = blockIdx.x; u < c.nunits; u += gridDim.x) {
    uint64_t tile_base;
    uint32_t item;
    if (c.units) {
      uint2 uu = c.units[u];
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
}

// ---------------------------------------------------------------------------
// Anchored leftmost-first walk: one lane per candidate start.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kTPB) void k_rx_walk(WalkParams w) {
  const uint32_t n = min(*w.ncand, w.cand_cap);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const Rec cd = w.cands[i];
    const uint32_t rx = cd.x, item = cd.z;
    const uint2 seg = w.items[item];
    const uint32_t* T = w.tables + w.toff[rx];
    const uint32_t C = w.ncls[rx];
    const uint8_t* cm = w.cmaps + 256u * rx;
    uint32_t s = 1, last = 0xFFFFFFFFu;
    uint64_t p = (uint64_t)seg.x + cd.y;
    bool dead = false;
    for (; p < seg.y; ++p) {
      const uint32_t e = T[s * C + cm[w.base[p]]];
      if (e & DFA_DEAD) { dead = true; break; }
      s = e & DFA_STATE_MASK;
      if (e & DFA_MATCH) last = (uint32_t)(p + 1 - seg.x);
    }
    if (!dead && w.window_truncated) atomicOr(&w.counters[3], kFlagTruncatedWalk);
    if (last != 0xFFFFFFFFu) {
      const uint32_t slot = atomicAdd(&w.counters[2], 1u);
      if (slot < w.raw_cap) w.raw[slot] = make_uint4(rx, cd.y, last, item);
    }
  }
}

// Chain keys: group (item, regex) in the high bits, start in the low bits.
__global__ void k_chain_keys(const Rec* raw, uint32_t n, uint32_t sbits, uint32_t rbits,
                             uint64_t* keys, uint32_t* vals) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const Rec r = raw[i];
    keys[i] = (((uint64_t)r.w << rbits | r.x) << sbits) | r.y;
    vals[i] = i;
  }
}

// A record starts a chain segment when it opens its group or when no earlier
// record of the group can overlap it (none within max length ends after it).
// group = key >> gshift (all one group when gshift >= 64); max length comes
// from maxlen[rec.x] or, when maxlen is null, from maxlen_all.
__device__ __forceinline__ uint64_t group_of(uint64_t k, uint32_t gshift) {
  return gshift >= 64 ? 0ull : (k >> gshift);
}

__global__ void k_chain_heads(const uint64_t* keys, const uint32_t* vals, const Rec* raw,
                              uint32_t n, uint32_t gshift, const uint32_t* maxlen,
                              uint32_t maxlen_all, uint8_t* head) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint64_t g = group_of(keys[i], gshift);
    const Rec ri = raw[vals[i]];
    bool h = (i == 0) || (group_of(keys[i - 1], gshift) != g);
    if (!h) {
      const uint32_t ml = maxlen ? maxlen[ri.x] : maxlen_all;
      if (ml != 0xFFFFFFFFu) {
        h = true;
        for (uint32_t j = i; j-- > 0;) {
          if (group_of(keys[j], gshift) != g) break;
          const Rec rj = raw[vals[j]];
          if ((uint64_t)rj.y + ml <= ri.y) break;
          if (rj.z > ri.y) { h = false; break; }
        }
      }
    }
    head[i] = h;
  }
}

// Greedy non-overlapping selection inside each segment (find_iter semantics):
// keep a record when it starts at or after the end of the last kept one.
// With `ties`, records sharing the last kept (start, end) are kept too
// (HashScanner reports every id of the shortest length at a position).
// Output pattern id is rx_pid[rec.x] when rx_pid is set, else rec.x.
__global__ void k_chain_walk(const uint64_t* keys, const uint32_t* vals, const Rec* raw,
                             uint32_t n, uint32_t gshift, const uint8_t* head,
                             const uint32_t* carried, uint32_t seg_off, const uint32_t* rx_pid,
                             uint32_t ties, Rec* out, uint32_t out_cap, uint32_t* out_count,
                             uint32_t* carried_out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (!head[i]) continue;
    const uint64_t g = group_of(keys[i], gshift);
    uint32_t last_end = 0, last_start = 0xFFFFFFFFu;
    if (carried) {
      const uint32_t cv = carried[raw[vals[i]].x];
      last_end = cv > seg_off ? cv - seg_off : 0u;
    }
    for (uint32_t j = i; j < n; ++j) {
      if (j > i && (head[j] || group_of(keys[j], gshift) != g)) break;
      const Rec r = raw[vals[j]];
      const bool tie = ties && r.y == last_start && r.z == last_end;
      if (r.y < last_end && !tie) continue;
      last_end = r.z;
      last_start = r.y;
      const uint32_t slot = atomicAdd(out_count, 1u);
      if (slot < out_cap) out[slot] = make_uint4(rx_pid ? rx_pid[r.x] : r.x, r.y, r.z, r.w);
      if (carried_out) atomicMax(&carried_out[r.x], seg_off + r.z);
    }
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

// Final ordering keys: (item, start, pattern_id); the end rides as value.
__global__ void k_order_keys(const Rec* recs, uint32_t n, uint32_t sbits, uint32_t pbits,
                             uint64_t* keys, uint32_t* vals) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const Rec r = recs[i];
    keys[i] = ((((uint64_t)r.w << sbits) | r.y) << pbits) | r.x;
    vals[i] = r.z;
  }
}

struct MatchOut { uint32_t pid, start, end; };
struct TaggedOut { uint64_t id; uint32_t pid, start, end, pad; };

__global__ void k_pack_matches(const uint64_t* keys, const uint32_t* vals, uint32_t n,
                               uint32_t pbits, uint32_t start_off, MatchOut* out) {
  const uint64_t pmask = (1ull << pbits) - 1;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint64_t k = keys[i];
    MatchOut m;
    m.pid = (uint32_t)(k & pmask);
    m.start = (uint32_t)(k >> pbits) + start_off;
    m.end = vals[i] + start_off;
    out[i] = m;
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

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host launch wrappers (the runtime in engine.cpp only calls these).
// ---------------------------------------------------------------------------
template <typename F>
static hipError_t allow_lds(F* fn, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <int K>
static hipError_t launch_hash_k(bool ci, uint32_t grid, size_t shmem, hipStream_t st,
                                const ScanCommon& c, const HashParams& h) {
  if (ci) {
    hipError_t e = allow_lds(k_hash_scan<K, true>, shmem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_hash_scan<K, true>), dim3(grid), dim3(kTPB), shmem, st, c, h);
  } else {
    hipError_t e = allow_lds(k_hash_scan<K, false>, shmem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_hash_scan<K, false>), dim3(grid), dim3(kTPB), shmem, st, c, h);
  }
  return hipGetLastError();
}

size_t hash_shmem_bytes(uint32_t nwords) { return nwords * 4u + 2u * kRecCap * sizeof(Rec); }

hipError_t launch_hash(uint32_t K, bool ci, uint32_t grid, hipStream_t st, const ScanCommon& c,
                       const HashParams& h) {
  const size_t shmem = hash_shmem_bytes(h.nwords);
  switch (K) {
    case 1: return launch_hash_k<1>(ci, grid, shmem, st, c, h);
    case 2: return launch_hash_k<2>(ci, grid, shmem, st, c, h);
    case 3: return launch_hash_k<3>(ci, grid, shmem, st, c, h);
    case 4: return launch_hash_k<4>(ci, grid, shmem, st, c, h);
    case 6: return launch_hash_k<6>(ci, grid, shmem, st, c, h);
    case 8: return launch_hash_k<8>(ci, grid, shmem, st, c, h);
    case 12: return launch_hash_k<12>(ci, grid, shmem, st, c, h);
    case 16: return launch_hash_k<16>(ci, grid, shmem, st, c, h);
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

static uint32_t grid_for(uint32_t n) {
  uint32_t g = (n + kTPB - 1) / kTPB;
  return g < 1 ? 1 : (g > 65535 ? 65535 : g);
}

hipError_t launch_walk(uint32_t ncand_host, hipStream_t st, const WalkParams& w) {
  hipLaunchKernelGGL(k_rx_walk, dim3(grid_for(ncand_host)), dim3(kTPB), 0, st, w);
  return hipGetLastError();
}

hipError_t launch_order_keys(const Rec* recs, uint32_t n, uint32_t sbits, uint32_t pbits,
                             uint64_t* keys, uint32_t* vals, hipStream_t st) {
  hipLaunchKernelGGL(k_order_keys, dim3(grid_for(n)), dim3(kTPB), 0, st, recs, n, sbits, pbits, keys, vals);
  return hipGetLastError();
}

hipError_t launch_chain_keys(const Rec* raw, uint32_t n, uint32_t sbits, uint32_t rbits,
                             uint64_t* keys, uint32_t* vals, hipStream_t st) {
  hipLaunchKernelGGL(k_chain_keys, dim3(grid_for(n)), dim3(kTPB), 0, st, raw, n, sbits, rbits, keys, vals);
  return hipGetLastError();
}

hipError_t launch_chain_heads(const uint64_t* keys, const uint32_t* vals, const Rec* raw, uint32_t n,
                              uint32_t gshift, const uint32_t* maxlen, uint32_t maxlen_all, uint8_t* head,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_chain_heads, dim3(grid_for(n)), dim3(kTPB), 0, st, keys, vals, raw, n, gshift, maxlen,
                     maxlen_all, head);
  return hipGetLastError();
}

hipError_t launch_chain_walk(const uint64_t* keys, const uint32_t* vals, const Rec* raw, uint32_t n,
                             uint32_t gshift, const uint8_t* head, const uint32_t* carried,
                             uint32_t seg_off, const uint32_t* rx_pid, uint32_t ties, Rec* out,
                             uint32_t out_cap, uint32_t* out_count, uint32_t* carried_out,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_chain_walk, dim3(grid_for(n)), dim3(kTPB), 0, st, keys, vals, raw, n, gshift, head,
                     carried, seg_off, rx_pid, ties, out, out_cap, out_count, carried_out);
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

hipError_t launch_pack_matches(const uint64_t* keys, const uint32_t* vals, uint32_t n, uint32_t pbits,
                               uint32_t start_off, void* out, hipStream_t st) {
  hipLaunchKernelGGL(k_pack_matches, dim3(grid_for(n)), dim3(kTPB), 0, st, keys, vals, n, pbits, start_off,
                     reinterpret_cast<MatchOut*>(out));
  return hipGetLastError();
}

hipError_t launch_pack_tagged(const uint64_t* keys, const uint32_t* vals, uint32_t n, uint32_t sbits,
                              uint32_t pbits, const uint64_t* ids, void* out, hipStream_t st) {
  hipLaunchKernelGGL(k_pack_tagged, dim3(grid_for(n)), dim3(kTPB), 0, st, keys, vals, n, sbits, pbits, ids,
                     reinterpret_cast<TaggedOut*>(out));
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

// Explicit instantiations used by the host runtime.
#define WS_HASH_INST(K)                                                          \
  template __global__ void k_hash_scan<K, false>(ScanCommon, HashParams);        \
  template __global__ void k_hash_scan<K, true>(ScanCommon, HashParams);
WS_HASH_INST(1)
WS_HASH_INST(2)
WS_HASH_INST(3)
WS_HASH_INST(4)
WS_HASH_INST(6)
WS_HASH_INST(8)
WS_HASH_INST(12)
WS_HASH_INST(16)
template __global__ void k_ac_scan<true, 64>(ScanCommon, AcParams);
template __global__ void k_ac_scan<false, 64>(ScanCommon, AcParams);

}  // namespace ws
