// HIP kernels of the scan path, written for gfx950 (CDNA4, wave64).
//
//  k_hash_filter rolling-hash prefilter (16-64 positions per lane, bloom
//                filter in LDS) with the verification of its hits in the
//                same kernel: per-wave LDS hit rings, 64 hits a round.
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

// Next work group of a persistent block: pulled from the device work queue
// when there is one, else a static stride. Thread 0 claims one group ahead
// (`ahead` lives in its registers), so the atomic's round trip overlaps the
// current group's loads instead of stalling the block at the barrier.
// Must be called by every thread of the block.
__device__ __forceinline__ uint32_t next_group(const ScanCommon& c, uint32_t cur, bool first, uint32_t* s_slot,
                                               uint32_t& ahead) {
  if (!c.queue) return first ? blockIdx.x : cur + gridDim.x;
  if (!first) __syncthreads();
  if (threadIdx.x == 0) {
    *s_slot = first ? atomicAdd(c.queue, 1u) : ahead;
    ahead = atomicAdd(c.queue, 1u);
  }
  __syncthreads();
  return *s_slot;
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
// Filter bits of the PL positions [p0, p0 + PL) whose window (PL + 16 bytes)
// is in d, as PL/32 mask words (one word, low PL bits, when PL < 32). Per
// position: one byte extract, two 24-bit multiply-adds (rolling hash), one
// LDS read of the bloom word and a two-bit test.
template <int PL>
constexpr int mask_words() { return PL > 32 ? PL / 32 : 1; }

template <int K, bool CI, int PL>
__device__ __forceinline__ void filter_lane(const HashParams& h, const uint32_t* bloom, uint32_t* d,
                                            uint32_t* out) {
  if (CI) {
#pragma unroll
    for (int j = 0; j < PL / 4 + 4; ++j) d[j] = fold_word(d[j]);
  }
  uint32_t H = byte_of(d, 0), mask = 0;
#pragma unroll
  for (int t = 1; t < K; ++t) H = mad24_s(H, h.B, byte_of(d, t));
#pragma unroll
  for (int i = 0; i < PL; ++i) {
    if (i > 0) H = mad24_s(H, h.B, mad24_s(byte_of(d, i - 1), h.NBK, byte_of(d, i + K - 1)));
    // Blocked bloom: one LDS word, two bits (H[4:0] and H[s2+4:s2]); v_bfe
    // reads its offset from the low five bits of the offset operand.
    const uint32_t w = bloom[__builtin_amdgcn_ubfe(H, h.wshift, h.wbits)];
    const uint32_t hit = __builtin_amdgcn_ubfe(w, H, 1) & __builtin_amdgcn_ubfe(w, H >> h.s2, 1);
    // v_alignbit shifts the mask right and brings bit 0 of `hit` in at the
    // top: after 32 positions, position i sits in bit i.
    mask = __builtin_amdgcn_alignbit(hit, mask, 1);
    if ((i & 31) == 31) out[i >> 5] = mask;
  }
  if (PL < 32) out[0] = mask >> (32 - PL);
}

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk2(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ u16x2 upk2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }

// K >= 3: the same filter over the 32-bit dual hash D (two 16-bit
// polynomial hashes rolled together by v_pk_mad_u16). The bloom word's byte
// address is D & amask, its two bits D[20:16] and D[28:24], which the shifts
// take straight from the upper half and top byte of D (SDWA operand
// selects). The window bytes are first spread into 16-bit pairs (one v_perm
// per two bytes), so a multiply-add takes its byte as either half of a pair
// (VOP3P operand selects): per position about half a perm, two packed
// multiply-adds, one address op, the LDS read, two shifts, an and and the
// mask update.
template <int K, bool CI, int PL>
__device__ __forceinline__ void filter_lane_dual(const HashParams& h, const uint32_t* bloom, uint32_t* d,
                                                 uint32_t* out) {
  constexpr int NW = PL / 4 + 4;
  if (CI) {
#pragma unroll
    for (int j = 0; j < NW; ++j) d[j] = fold_word(d[j]);
  }
  // pr[2m] = {byte 4m, byte 4m+1}, pr[2m+1] = {byte 4m+2, byte 4m+3}.
  u16x2 pr[2 * NW];
#pragma unroll
  for (int m = 0; m < NW; ++m) {
    pr[2 * m] = upk2(__builtin_amdgcn_perm(0u, d[m], 0x0c010c00u));
    pr[2 * m + 1] = upk2(__builtin_amdgcn_perm(0u, d[m], 0x0c030c02u));
  }
  auto at = [&](int t) -> u16x2 {
    const u16x2 v = pr[t >> 1];
    return (t & 1) ? __builtin_shufflevector(v, v, 1, 1) : __builtin_shufflevector(v, v, 0, 0);
  };
  const u16x2 bb = upk2(h.BB), nbk = upk2(h.NBKK);
  u16x2 D = at(0);
  uint32_t mask = 0;
#pragma unroll
  for (int t = 1; t < K; ++t) D = D * bb + at(t);
  // The bloom words sit at LDS address 0 (k_hash_filter has no static LDS),
  // so the masked hash is the ds_read address as it stands.
  (void)bloom;
  const __attribute__((address_space(3))) char* lds = nullptr;
#pragma unroll
  for (int i = 0; i < PL; ++i) {
    if (i > 0) D = D * bb + (at(i - 1) * nbk + at(i + K - 1));
    const uint32_t Dw = pk2(D);
    const uint32_t w = *reinterpret_cast<const __attribute__((address_space(3))) uint32_t*>(lds + (Dw & h.amask));
    const uint32_t hit = (w >> ((Dw >> 16) & 31u)) & (w >> ((Dw >> 24) & 31u));
    mask = __builtin_amdgcn_alignbit(hit, mask, 1);
    if ((i & 31) == 31) out[i >> 5] = mask;
  }
  if (PL < 32) out[0] = mask >> (32 - PL);
}

// Valid positions q of [p0, p0 + n) (n <= 32): lo <= q, q + K <= hi,
// q - lo < emit_end.
__device__ __forceinline__ uint32_t valid_word(uint32_t p0, uint32_t n, uint32_t lo, uint32_t hi, uint32_t K,
                                               uint32_t emit_end) {
  if (hi < K) return 0;
  uint32_t qmin = p0 > lo ? p0 : lo;
  uint32_t qmax = hi - K + 1;
  if (emit_end < qmax - lo) qmax = lo + emit_end;
  if (qmax > p0 + n) qmax = p0 + n;
  if (qmax <= qmin) return 0;
  const uint32_t a0 = qmin - p0, a1 = qmax - p0;
  const uint32_t top = a1 >= 32 ? 0xFFFFFFFFu : ((1u << a1) - 1u);
  return top & ~((1u << a0) - 1u);
}

constexpr int kUnitsPerIter = 2;  // quarter units a lane hashes per iteration (a wave: one half unit)

__device__ __forceinline__ void slice_put(const SliceOut& so, uint32_t slice, uint32_t slot, Rec r) {
  if (slot < so.cap) {
    so.buf[(size_t)slice * so.cap + slot] = r;
  } else {
    const uint32_t g = atomicAdd(so.ovf_count, 1u);
    if (g < so.ovf_cap) so.ovf[g] = r;
  }
}

// The 32 bytes from q as eight dwords, out of the 48-byte aligned span
// raw[0..11] that holds them (q - (q & ~15) = sh): two operand selects by
// the dword shift, then one v_alignbyte per output dword.
// (Bit selects through all-ones / zero masks, v_bfi_b32, so the compiler
// keeps the span in registers instead of indexing a stack copy.)
__device__ __forceinline__ void realign32(const uint32_t* raw, uint32_t sh, uint32_t* W) {
  const uint32_t m2 = 0u - ((sh >> 3) & 1u), m1 = 0u - ((sh >> 2) & 1u);
  uint32_t t[10], u[9];
#pragma unroll
  for (int i = 0; i < 10; ++i) t[i] = (raw[i + 2] & m2) | (raw[i] & ~m2);
#pragma unroll
  for (int i = 0; i < 9; ++i) u[i] = (t[i + 1] & m1) | (t[i] & ~m1);
#pragma unroll
  for (int i = 0; i < 8; ++i) W[i] = __builtin_amdgcn_alignbyte(u[i + 1], u[i], sh & 3);
}

// Exact comparison of one table entry ({hash, factor byte offset (16-byte
// aligned), length | kind << 31, bytes 0-3}, {bytes 4-15, output id}) at
// position q. W holds the (folded, when CI) input bytes q..q+31. On a match
// *rec is the output record and the result is 1 (literal) or 2 (regex
// candidate); 0 otherwise.
template <bool CI>
__device__ __forceinline__ uint32_t verify_entry(const ScanCommon& c, uint4 e0, uint4 e1, const uint32_t* W,
                                                 uint64_t q, uint2 seg, uint32_t item, uint4* rec) {
  const uint32_t len = e0.z & 0x7FFFFFFFu;
  if (q + len > seg.y) return 0;
  const uint32_t pre[4] = {e0.w, e1.x, e1.y, e1.z};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (4u * k >= len) break;
    const uint32_t rem = len - 4 * k;
    const uint32_t msk = rem >= 4 ? 0xFFFFFFFFu : ((1u << (8 * rem)) - 1u);
    if ((W[k] ^ pre[k]) & msk) return 0;
  }
  const uint32_t start = (uint32_t)(q - seg.x);
  const uint32_t kind = (e0.z >> 31) ? 2u : 1u;
  *rec = make_uint4(e1.w, start, kind == 2u ? 0u : start + len, item);
  if (len > 16) {
    // Bytes 16..31 of the factor in one load.
    const uint4 p1 = *reinterpret_cast<const uint4*>(c.fbytes + e0.y + 16);
    const uint32_t pw[4] = {p1.x, p1.y, p1.z, p1.w};
    uint32_t diff = 0;
#pragma unroll
    for (int k = 4; k < 8; ++k) {
      const uint32_t rem = len > 4u * k ? len - 4 * k : 0u;
      const uint32_t msk = rem >= 4 ? 0xFFFFFFFFu : ((1u << (8 * rem)) - 1u);
      diff |= (W[k] ^ pw[k - 4]) & msk;
    }
    if (diff) return 0;
    // Factors longer than 32 bytes: the rest byte by byte.
    const uint32_t nw = (len + 3) >> 2;
    const uint32_t* pat = reinterpret_cast<const uint32_t*>(c.fbytes + e0.y);
    for (uint32_t k = 8; k < nw; ++k) {
      uint32_t dv = 0;
      for (int t = 0; t < 4; ++t) {
        const uint64_t p = q + 4 * k + t;
        uint8_t bb = p < seg.y ? c.base[p] : 0;
        if (CI) bb = fold_byte(bb);
        dv |= (uint32_t)bb << (8 * t);
      }
      const uint32_t rem = len - 4 * k;
      const uint32_t msk = rem >= 4 ? 0xFFFFFFFFu : ((1u << (8 * rem)) - 1u);
      if ((dv ^ pat[k]) & msk) return 0;
    }
  }
  return kind;
}

// Appends rec for every lane with `take` to the block's slice of `so`: one
// LDS atomic per wave (ballot + prefix count). Call with every active lane.
__device__ __forceinline__ void emit_wave(const SliceOut& so, uint32_t* s_cnt, bool take, uint4 rec) {
  const uint64_t m = __ballot(take);
  if (m == 0) return;
  const uint32_t lane = __lane_id();
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if ((int)lane == leader) base = atomicAdd(s_cnt, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  if (take) slice_put(so, blockIdx.x, base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)), rec);
}

// A filter hit being verified: its position, item, segment, the input
// bytes from q, and the window hash.
struct HitWork {
  uint64_t q;
  uint32_t item;
  uint2 seg;
  uint32_t W[8];
  uint32_t H;
  uint4 s0, s1, s2, s3;  // its bucket's two slots
  bool live;             // false for a lane without a hit this round
};

// raw: the aligned span from q & ~15, its first 32 bytes loaded (zeros past
// the loadable end). The window hash is checked against the bucket's two
// tags first; only a match (or a crowded bucket) reads the slots and the
// span's third 16 bytes (for factors that run past the first 32).
template <int K, bool CI>
__device__ __forceinline__ void hit_prepare(const ScanCommon& c, const HashParams& h, uint2 hr, uint2 seg,
                                            uint32_t* raw, uint32_t lend, HitWork& hw) {
  hw.live = true;
  hw.q = hr.x;
  hw.item = hr.y;
  hw.seg = seg;
  const uint32_t sh = (uint32_t)(hw.q & 15);
  realign32(raw, sh, hw.W);
  // The verification table's 24-bit window hash (v_mad_u32_u24 reads the
  // low 24 bits of the running value; the window lies in the first 16 bytes).
  uint32_t H = 0;
#pragma unroll
  for (int t = 0; t < K; ++t) {
    const uint32_t b = (hw.W[t >> 2] >> ((t & 3) * 8)) & 0xFFu;
    H = mad24_s(H, h.B, CI ? fold_byte((uint8_t)b) : b);
  }
  H &= 0xFFFFFFu;
  hw.H = H;
  const size_t bucket = (H * 0x9E3779B1u) >> (32 - h.nbits);
  const uint2 tg = h.vtag[bucket];
  hw.s0.x = kVtabEmpty;
  hw.s2.x = kVtabEmpty;
  const bool m0 = tg.x == H, m1 = tg.y == H || tg.y == kVtabMore;
  if (m0 || m1) {
    const uint32_t a = (uint32_t)(hw.q & ~15ull) + 32;
    if (a + 16 <= lend) {
      const uint4 x2 = *reinterpret_cast<const uint4*>(c.base + a);
      raw[8] = x2.x; raw[9] = x2.y; raw[10] = x2.z; raw[11] = x2.w;
      realign32(raw, sh, hw.W);
    }
    const uint4* slot = h.vtab + 4 * bucket;
    if (m0) {
      hw.s0 = slot[0];
      hw.s1 = slot[1];
    }
    if (m1) {
      hw.s2 = slot[2];
      hw.s3 = slot[3];
    }
  }
  if (CI) {
#pragma unroll
    for (int k = 0; k < 8; ++k) hw.W[k] = fold_word(hw.W[k]);
  }
}

// RX: the set has regex factors (kind-2 entries, candidate output).
template <bool CI, bool RX>
__device__ __forceinline__ void hit_finish(const ScanCommon& c, const HashParams& h, const HitWork& hw,
                                           uint32_t* s_n, const SliceOut& lit, const SliceOut& cand) {
  uint4 r0 = make_uint4(0, 0, 0, 0), r1 = r0;
  uint32_t k0 = 0, k1 = 0;
  if (hw.live && hw.s0.x == hw.H)
    k0 = verify_entry<CI>(c, hw.s0, hw.s1, hw.W, hw.q, hw.seg, hw.item, &r0);
  if (!hw.live) {
  } else if (hw.s2.x == hw.H) {
    k1 = verify_entry<CI>(c, hw.s2, hw.s3, hw.W, hw.q, hw.seg, hw.item, &r1);
  } else if (hw.s2.x == kVtabMore) {
    // Crowded bucket (rare): per-lane appends.
    for (uint32_t e = hw.s2.y; e < hw.s2.y + hw.s2.z; ++e) {
      const uint4 e0 = h.vmore[2 * (size_t)e];
      if (e0.x != hw.H) continue;
      uint4 rec;
      const uint32_t k = verify_entry<CI>(c, e0, h.vmore[2 * (size_t)e + 1], hw.W, hw.q, hw.seg, hw.item, &rec);
      if (k == 1) slice_put(lit, blockIdx.x, atomicAdd(&s_n[0], 1u), rec);
      if (RX && k == 2) slice_put(cand, blockIdx.x, atomicAdd(&s_n[1], 1u), rec);
    }
  }
  emit_wave(lit, &s_n[0], k0 == 1, r0);
  if constexpr (RX) emit_wave(cand, &s_n[1], k0 == 2, r0);
  emit_wave(lit, &s_n[0], k1 == 1, r1);
  if constexpr (RX) emit_wave(cand, &s_n[1], k1 == 2, r1);
}

// Per-wave LDS ring of filter hits waiting for verification (uint2
// {position, item}); a wave verifies 64 at a time, one per lane.
constexpr uint32_t kHitRing = 128;

// Verifies the n (<= 64) queued hits from ring index `head`, one per lane:
// the 32-byte aligned input span holding the window (streamed by the filter
// shortly before), the window hash, its bucket's tags, and for a tag match
// the slots and the exact compare.
// Call with the whole wave.
// seg0: the segment of a single-segment scan (c.units null), wave-uniform.
template <int K, bool CI, bool RX>
__device__ __forceinline__ void verify_queued(const ScanCommon& c, const HashParams& h, const uint2* ring,
                                              uint32_t head, uint32_t n, uint32_t* s_cnt, const SliceOut& lit,
                                              const SliceOut& cand, uint2 seg0) {
  const uint32_t lane = lane_id();
  HitWork hw{};
  hw.live = false;
  if (lane < n) {
    const uint2 hr = ring[(head + lane) & (kHitRing - 1)];
    const uint2 seg = c.units ? c.items[hr.y] : seg0;
    uint32_t lend = (seg.y + 15) & ~15u;
    if (lend > (uint32_t)c.load_end) lend = (uint32_t)c.load_end;
    const uint32_t a = hr.x & ~15u;
    uint32_t raw[12];
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      uint4 x = make_uint4(0, 0, 0, 0);
      if (a + 16 * v + 16 <= lend) x = *reinterpret_cast<const uint4*>(c.base + a + 16 * v);
      raw[4 * v] = x.x; raw[4 * v + 1] = x.y; raw[4 * v + 2] = x.z; raw[4 * v + 3] = x.w;
    }
    raw[8] = raw[9] = raw[10] = raw[11] = 0;
    hit_prepare<K, CI>(c, h, hr, seg, raw, lend, hw);
  }
  hit_finish<CI, RX>(c, h, hw, s_cnt, lit, cand);
}

// Kernel duration for profiling without events (which a replayed graph does
// not re-record): the first block in takes the minimum start, the last one
// out the maximum end, on the device's constant-rate wall clock.
__device__ __forceinline__ void stamp_start(const ScanCommon& c) {
  if (c.stamp && threadIdx.x == 0) atomicMin(&c.stamp[0], (unsigned long long)wall_clock64());
}
__device__ __forceinline__ void stamp_end(const ScanCommon& c) {
  if (c.stamp && threadIdx.x == 0) atomicMax(&c.stamp[1], (unsigned long long)wall_clock64());
}

// Orders a wave's LDS ring writes before its lanes read each other's entries.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Streaming rolling-hash filter with in-kernel verification, one 1024-lane
// block per CU so the bloom filter can use up to 128 KiB of LDS. Work goes
// out in half units (a unit is 256 * PL bytes; the host lists a batch's
// units), 16 to a group, the groups striding over the blocks; a block's
// waves take its half units one at a time from an LDS counter, so the waves
// the scheduler serves more often do more and all finish together. (A
// device-wide counter was measured slower: its atomics cost more than the
// imbalance across CUs they remove.) Per half unit a lane loads the
// (PL + 16)-byte windows of its two quarters with coalesced 16-byte loads,
// then hashes PL positions in each. Tiles wholly inside their
// segment skip all bounds work. A wave appends its hits (wave prefix sum) to
// its own LDS ring and verifies them 64 at a time while the block's other
// waves keep filtering, so the verification's dependent L2 reads hide
// behind their arithmetic: no hit list goes through HBM and the loop has no
// block barriers. Literal matches and regex candidate starts go to the
// block's output slices.
// The filter's 32-bit bounds at the largest span (16-byte shift included):
// the rounded segment end and the last half unit's end stay below 2^32.
static_assert(kMaxScanBytes + 16 + 15 < (1ull << 32), "segment end wraps");
static_assert(kMaxScanBytes + 16 + (uint64_t)kTPB * 64 + 16 < (1ull << 32), "half-unit bound wraps");

template <int K, bool CI, int PL, bool RX>
__global__ __launch_bounds__(kFilterTPB) void k_hash_filter(ScanCommon c, HashParams h, SliceOut lit,
                                                             SliceOut cand) {
  // LDS: the bloom words from address 0 (the kernel has no static LDS, so
  // a word's address is the masked hash itself), the block's two output
  // counters, its work counter and its hit count, then one hit ring per wave.
  extern __shared__ uint32_t smem[];
  uint32_t* bloom = smem;
  uint32_t* s_cnt = &smem[h.nwords];
  uint32_t* s_work = &smem[h.nwords + 2];
  uint32_t* s_hits = &smem[h.nwords + 3];
  uint2* ring = reinterpret_cast<uint2*>(&smem[h.nwords + 4]) + (threadIdx.x / 64) * kHitRing;
  constexpr int NW = PL / 4 + 4;  // window dwords
  constexpr uint32_t kUnit = kTPB * PL;
  constexpr uint32_t kWaves = kFilterTPB / 64;
  const uint32_t lane = lane_id();
  stamp_start(c);

  {
    // The bloom filter (nwords is a power of two >= 8) in 16-byte pieces;
    // whole passes of eight loads per lane put them all in flight before
    // any store.
    const uint4* src = reinterpret_cast<const uint4*>(h.bloom);
    uint4* dst = reinterpret_cast<uint4*>(bloom);
    const uint32_t n4 = h.nwords / 4;
    if (n4 % (8 * kFilterTPB) == 0) {
      constexpr uint32_t T = kFilterTPB;
      for (uint32_t k0 = threadIdx.x; k0 < n4; k0 += 8 * T) {
        const uint4 a0 = src[k0], a1 = src[k0 + T], a2 = src[k0 + 2 * T], a3 = src[k0 + 3 * T];
        const uint4 a4 = src[k0 + 4 * T], a5 = src[k0 + 5 * T], a6 = src[k0 + 6 * T], a7 = src[k0 + 7 * T];
        dst[k0] = a0; dst[k0 + T] = a1; dst[k0 + 2 * T] = a2; dst[k0 + 3 * T] = a3;
        dst[k0 + 4 * T] = a4; dst[k0 + 5 * T] = a5; dst[k0 + 6 * T] = a6; dst[k0 + 7 * T] = a7;
      }
    } else {
      for (uint32_t k = threadIdx.x; k < n4; k += kFilterTPB) dst[k] = src[k];
    }
  }
  if (threadIdx.x == 0) {
    s_cnt[0] = 0;
    s_cnt[1] = 0;
    *s_work = 0;
    *s_hits = 0;
  }
  __syncthreads();

  const uint32_t load_end = (uint32_t)c.load_end;
  const uint32_t nslots = 2 * c.nunits;
  // The block's jw-th half unit: slot jw % 16 of group blockIdx.x +
  // (jw / 16) * gridDim.x, increasing in jw, so the first one past the end
  // means the block is done.
  auto claim = [&]() -> uint32_t {
    uint32_t v = 0;
    if (lane == 0) v = atomicAdd(s_work, 1u);
    const uint32_t jw = __builtin_amdgcn_readfirstlane(v);
    return (blockIdx.x + (jw / kWaves) * gridDim.x) * kWaves + jw % kWaves;
  };
  // A half unit's bounds (wave-uniform): from its unit's record in a
  // batch, else from the one segment, read once.
  struct Half {
    uint32_t s, tile_base, item, lo, hi, lend;
    uint32_t inner;  // every window byte inside the segment, every position may start a match
  };
  const uint2 seg0 = c.units ? make_uint2(0, 0) : c.items[0];
  auto half_info = [&](uint32_t sv, uint4 rec) {
    Half hf;
    hf.s = sv;
    if (c.units) {
      hf.tile_base = rec.x;
      hf.item = rec.y;
      hf.lo = rec.z;
      hf.hi = rec.w;
    } else {
      hf.tile_base = (sv >> 1) * kUnit;
      hf.item = 0;
      hf.lo = seg0.x;
      hf.hi = seg0.y;
    }
    hf.inner = hf.tile_base >= hf.lo && hf.tile_base + kUnit + 16 <= hf.hi &&
               hf.tile_base + kUnit - hf.lo + 0u <= c.emit_end;
    hf.lend = (hf.hi + 15) & ~15u;
    if (hf.lend > load_end) hf.lend = load_end;
    return hf;
  };
  // A batch unit's record, loaded one half unit ahead of its use.
  auto unit_rec = [&](uint32_t sv) { return c.units && sv < nslots ? c.units[sv >> 1] : make_uint4(0, 0, 0, 0); };
  // Quarter 2 * (s & 1) + j of unit s / 2: 64 lanes x PL bytes.
  auto quarter_p0 = [&](const Half& hf, int j) { return hf.tile_base + PL * (64 * (2 * (hf.s & 1) + j) + lane); };
  // Granules outside the segment are clamped to its first or last one: one
  // load per granule on every path (so the waits before each use can count
  // them), and positions whose window leaves the segment are masked out
  // afterwards, whatever bytes they hashed.
  auto load_quarter = [&](const Half& hf, int j, uint32_t* dj) {
    const uint32_t p0 = quarter_p0(hf, j);
    const uint32_t amin = hf.lo & ~15u, amax = hf.lend - 16;
#pragma unroll
    for (int v = 0; v < NW / 4; ++v) {
      uint32_t a = p0 + 16 * v;
      a = a < amin ? amin : (a > amax ? amax : a);
      const uint4 x = *reinterpret_cast<const uint4*>(c.base + a);
      dj[4 * v] = x.x; dj[4 * v + 1] = x.y; dj[4 * v + 2] = x.z; dj[4 * v + 3] = x.w;
    }
  };
  constexpr int NM = mask_words<PL>();
  constexpr uint32_t WPOS = PL < 32 ? PL : 32;  // positions per mask word
  auto filter_quarter = [&](const Half& hf, int j, uint32_t* dj, uint32_t* mj) {
    const uint32_t p0 = quarter_p0(hf, j);
#pragma unroll
    for (int k = 0; k < NM; ++k) mj[k] = 0;
    if (hf.inner || (p0 < hf.hi && p0 + PL > hf.lo)) {
      if constexpr (K >= 3)
        filter_lane_dual<K, CI, PL>(h, bloom, dj, mj);
      else
        filter_lane<K, CI, PL>(h, bloom, dj, mj);
      if (!hf.inner) {
#pragma unroll
        for (int k = 0; k < NM; ++k) mj[k] &= valid_word(p0 + WPOS * k, WPOS, hf.lo, hf.hi, K, c.emit_end);
      }
    }
  };

  // Software pipeline over the wave's half units: a quarter's window is
  // reloaded with the next half unit's as soon as it has been hashed, so
  // one quarter's loads are always in flight behind the arithmetic. (After
  // the last half unit the loads repeat it: every path issues the same
  // loads, so the waits before each use can count them.)
  uint32_t tail = 0, pend = 0, nhits = 0;  // the wave's ring (same in every lane)
  uint32_t d[kUnitsPerIter][NW];
  uint32_t cur_s = claim(), nxt_s = nslots;
  Half cur{};
  uint4 nxt_rec = make_uint4(0, 0, 0, 0);
  if (cur_s < nslots) {
    cur = half_info(cur_s, unit_rec(cur_s));
    load_quarter(cur, 0, d[0]);
    load_quarter(cur, 1, d[1]);
    nxt_s = claim();
    nxt_rec = unit_rec(nxt_s);
  }
  while (cur_s < nslots) {
    const bool more = nxt_s < nslots;
    const Half nxt = more ? half_info(nxt_s, nxt_rec) : cur;
    uint32_t mask[kUnitsPerIter][NM];
    filter_quarter(cur, 0, d[0], mask[0]);
    load_quarter(nxt, 0, d[0]);
    filter_quarter(cur, 1, d[1], mask[1]);
    load_quarter(nxt, 1, d[1]);
    uint32_t p0[kUnitsPerIter], item[kUnitsPerIter];
    uint32_t nh = 0;
#pragma unroll
    for (int j = 0; j < kUnitsPerIter; ++j) {
      p0[j] = quarter_p0(cur, j);
      item[j] = cur.item;
#pragma unroll
      for (int k = 0; k < NM; ++k) nh += __builtin_popcount(mask[j][k]);
    }
    cur_s = nxt_s;
    cur = nxt;
    if (more) {
      nxt_s = claim();
      nxt_rec = unit_rec(nxt_s);
    }
    if (__any(nh != 0u)) {
      // Rank of the lane's first hit among the wave's T hits.
      uint32_t x = nh;
#pragma unroll
      for (int dd = 1; dd < 64; dd <<= 1) {
        const uint32_t y = __shfl_up(x, dd, 64);
        if (lane >= (uint32_t)dd) x += y;
      }
      const uint32_t T = __shfl(x, 63, 64), off = x - nh;
      nhits += T;
      // Ranks [lo_r, lo_r + take) go to the ring per pass (one pass unless
      // the ring is short of room), each pass followed by full rounds.
      for (uint32_t lo_r = 0; lo_r < T;) {
        const uint32_t take = min(kHitRing - pend, T - lo_r);
        if (off < lo_r + take && off + nh > lo_r) {
          uint32_t r = off;
#pragma unroll
          for (int j = 0; j < kUnitsPerIter; ++j) {
#pragma unroll
            for (int k = 0; k < NM; ++k)
              for (uint32_t m = mask[j][k]; m; m &= m - 1, ++r)
                if (r - lo_r < take)
                  ring[(tail + r - lo_r) & (kHitRing - 1)] = make_uint2(p0[j] + WPOS * k + __builtin_ctz(m), item[j]);
          }
        }
        tail += take;
        pend += take;
        lo_r += take;
        wave_lds_sync();
        while (pend >= 64) {
          verify_queued<K, CI, RX>(c, h, ring, tail - pend, 64, s_cnt, lit, cand, seg0);
          pend -= 64;
        }
        wave_lds_sync();
      }
    }
  }
  while (pend > 0) {
    const uint32_t n = min(pend, 64u);
    verify_queued<K, CI, RX>(c, h, ring, tail - pend, n, s_cnt, lit, cand, seg0);
    pend -= n;
  }
  // Filter hits (for timings): one device atomic per block, not per wave;
  // same-address atomics from thousands of waves serialise at the end of a
  // short launch.
  if (lane == 0 && nhits) atomicAdd(s_hits, nhits);
  __syncthreads();
  if (threadIdx.x == 0) {
    lit.counts[blockIdx.x] = s_cnt[0];
    cand.counts[blockIdx.x] = s_cnt[1];
    if (*s_hits) atomicAdd(&c.counters[8], *s_hits);
  }
  stamp_end(c);
}

// Slice compaction of both verify outputs. One 1024-lane block: lanes
// [0, 512) take an exclusive scan of min(count, cap) over the literal
// slices, lanes [512, 1024) over the candidate slices; offsets[n] is the
// slices' sum, and the total (overflow area included) lands in *lit_total /
// *cand_total, the literal total also in *lit_snap when set.
__global__ __launch_bounds__(1024) void k_slice_offsets2(SliceOut lit, SliceOut cand, uint32_t n, uint32_t* loffs,
                                                          uint32_t* coffs, uint32_t* lit_total, uint32_t* cand_total,
                                                          uint32_t* lit_snap) {
  __shared__ uint32_t part[1024];
  const uint32_t half = threadIdx.x >> 9, t = threadIdx.x & 511;
  const SliceOut& so = half ? cand : lit;
  uint32_t* offsets = half ? coffs : loffs;
  const uint32_t per = (n + 511) / 512;
  const uint32_t b0 = t * per;
  uint32_t sum = 0;
  for (uint32_t k = 0; k < per && b0 + k < n; ++k) sum += min(so.counts[b0 + k], so.cap);
  part[threadIdx.x] = sum;
  __syncthreads();
  for (uint32_t d = 1; d < 512; d <<= 1) {
    uint32_t v = t >= d ? part[threadIdx.x - d] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - sum;
  for (uint32_t k = 0; k < per && b0 + k < n; ++k) {
    offsets[b0 + k] = run;
    run += min(so.counts[b0 + k], so.cap);
  }
  if (t == 511) {
    offsets[n] = part[threadIdx.x];
    const uint32_t tot = part[threadIdx.x] + min(*so.ovf_count, so.ovf_cap);
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
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
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
  if (threadIdx.x == 0) s_res[0] = 0;
  __syncthreads();
  const uint32_t local = want ? atomicAdd(&s_res[0], want) : 0u;
  __syncthreads();
  if (threadIdx.x == 0) s_res[1] = s_res[0] ? atomicAdd(counter, s_res[0]) : 0u;
  __syncthreads();
  return s_res[1] + local;
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
__device__ __forceinline__ uint32_t rx_match_at(const WalkParams& w, uint32_t rx, uint2 seg, uint32_t st,
                                                bool* open) {
  const uint32_t* T = w.tables + w.toff[rx];
  const uint32_t C = w.ncls[rx];
  const uint8_t* cm = w.cmaps + 256u * rx;
  const uint64_t p0 = (uint64_t)seg.x + st;
  const uint32_t ctx = (st > 0 || w.ctx_before) ? byte_ctx(w.base[p0 - 1]) : 0u;
  uint32_t s = w.starts[4 * rx + ctx];
  uint32_t last = 0xFFFFFFFFu;
  for (uint64_t p = p0; p < seg.y; ++p) {
    const uint32_t e = T[s * C + cm[w.base[p]]];
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
      last = rx_match_at(w, cd.x, w.items[cd.w], cd.y, &open);
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

// With carry_dst set, the kernel also copies the ncarry words of carry_src
// there (whatever the count): the chain walk that follows reads carry_src,
// the previous chunk's final value, and raises carry_dst, so no lane sees a
// value another lane of the same launch has already raised.
__global__ void k_chain_heads(const uint64_t* keys, const uint32_t* vals, const Rec* raw,
                              uint32_t n, const uint32_t* nd, uint32_t gshift, const uint32_t* maxlen,
                              uint32_t maxlen_all, uint8_t* head, const uint32_t* carry_src, uint32_t* carry_dst,
                              uint32_t ncarry) {
  if (carry_dst)
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < ncarry; k += gridDim.x * blockDim.x)
      carry_dst[k] = carry_src[k];
  n = count_of(n, nd);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint64_t g = group_at(keys, vals, raw, i, gshift);
    const Rec ri = raw[vals[i]];
    bool h = (i == 0) || (group_at(keys, vals, raw, i - 1, gshift) != g);
    if (!h) {
      // An empty record also depends on an earlier end equal to its start
      // (find_iter skips an empty match where the previous one ended).
      const bool empty = ri.y == ri.z;
      const uint32_t ml = maxlen ? maxlen[ri.x] : maxlen_all;
      if (ml != 0xFFFFFFFFu) {
        h = true;
        for (uint32_t j = i; j-- > 0;) {
          if (group_at(keys, vals, raw, j, gshift) != g) break;
          const Rec rj = raw[vals[j]];
          const uint64_t reach = (uint64_t)rj.y + ml;
          if (reach < ri.y || (reach == ri.y && !empty)) break;
          if (rj.z > ri.y || (empty && rj.z == ri.y)) { h = false; break; }
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
// A head lane walks its chain twice: once to count the records it keeps,
// then, after the block has reserved all its slots at once, to write them.
// Stream chunks: carried[rx] is 1 + the absolute end of the regex's last
// kept match in earlier chunks (0: none), read-only in this launch;
// carried_out (a different buffer, holding a copy of carried) is raised to
// this launch's kept ends. Both passes of a lane therefore see the same
// carried value.
// Launched with kTPB threads per block.
__global__ __launch_bounds__(kTPB) void k_chain_walk(const uint64_t* keys, const uint32_t* vals, const Rec* raw,
                                                     uint32_t n, const uint32_t* nd, uint32_t gshift,
                                                     const uint8_t* head, const uint32_t* carried, uint32_t seg_off,
                                                     const uint32_t* rx_pid, uint32_t ties, Rec* out,
                                                     uint32_t out_cap, uint32_t* out_count, uint32_t* carried_out,
                                                     OrderOut oo) {
  __shared__ uint32_t s_res[2];
  n = count_of(n, nd);
  // The chain from head i: with `emit`, its kept records go to slots
  // slot, slot + 1, ...; returns how many it keeps.
  auto walk = [&](uint32_t i, bool emit, uint32_t slot) -> uint32_t {
    const uint64_t g = group_at(keys, vals, raw, i, gshift);
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
      if (j > i && (head[j] || group_at(keys, vals, raw, j, gshift) != g)) break;
      const Rec r = raw[vals[j]];
      const bool tie = ties && r.y == last_start && r.z == last_end;
      if (r.y < last_end && !tie) continue;
      // find_iter: an empty match where the previous match ended is skipped.
      if (r.y == r.z && have_last && r.y == last_end && !tie) continue;
      last_end = r.z;
      last_start = r.y;
      have_last = true;
      if (emit) {
        const uint32_t sl = slot + kept;
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
      }
      ++kept;
    }
    return kept;
  };
  // Block-uniform loop: every lane reaches the slot reservation.
  for (uint32_t b0 = blockIdx.x * kTPB; b0 < n; b0 += gridDim.x * kTPB) {
    const uint32_t i = b0 + threadIdx.x;
    const bool on = i < n && head[i];
    const uint32_t kept = on ? walk(i, false, 0) : 0u;
    const uint32_t slot = block_reserve(kept, out_count, s_res);
    if (kept) walk(i, true, slot);
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
// Raises a kernel's dynamic LDS limit once (remembered per kernel), so
// launches inside a graph capture make no attribute calls.
static std::mutex g_lds_mu;
static std::map<const void*, size_t> g_lds_set;

template <typename F>
static hipError_t allow_lds(F* fn, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  const void* key = reinterpret_cast<const void*>(fn);
  std::lock_guard<std::mutex> lk(g_lds_mu);
  auto it = g_lds_set.find(key);
  if (it != g_lds_set.end() && it->second >= bytes) return hipSuccess;
  hipError_t e = hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) g_lds_set[key] = bytes;
  return e;
}

size_t hash_shmem_bytes(uint32_t nwords) { return nwords * 4u + 16u + (kFilterTPB / 64) * kHitRing * sizeof(uint2); }

template <int K, bool CI, bool RX>
static hipError_t prepare_hash_kcr(uint32_t lane_bytes, size_t shmem) {
  if (lane_bytes == 64) return allow_lds(k_hash_filter<K, CI, 64, RX>, shmem);
  if (lane_bytes == 32) return allow_lds(k_hash_filter<K, CI, 32, RX>, shmem);
  return allow_lds(k_hash_filter<K, CI, 16, RX>, shmem);
}

template <int K, bool CI>
static hipError_t prepare_hash_kc(uint32_t lane_bytes, size_t shmem, bool rx) {
  return rx ? prepare_hash_kcr<K, CI, true>(lane_bytes, shmem) : prepare_hash_kcr<K, CI, false>(lane_bytes, shmem);
}

// Sets the filter kernel's dynamic LDS limit ahead of launches that may be
// captured into a graph.
hipError_t prepare_hash(uint32_t K, bool ci, uint32_t lane_bytes, uint32_t nwords, bool rx) {
  const size_t shmem = hash_shmem_bytes(nwords);
  switch (K) {
#define WS_K(k) \
  case k: return ci ? prepare_hash_kc<k, true>(lane_bytes, shmem, rx) : prepare_hash_kc<k, false>(lane_bytes, shmem, rx);
    WS_K(1) WS_K(2) WS_K(3) WS_K(4) WS_K(6) WS_K(8) WS_K(12) WS_K(16)
#undef WS_K
    default: return hipErrorInvalidValue;
  }
}

template <int K, bool CI, bool RX>
static hipError_t launch_hash_kcr(uint32_t grid, hipStream_t st, const ScanCommon& c, const HashParams& h,
                                  const SliceOut& lit, const SliceOut& cand) {
  const size_t shmem = hash_shmem_bytes(h.nwords);
  if (h.lane_bytes == 64) {
    hipError_t e = allow_lds(k_hash_filter<K, CI, 64, RX>, shmem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_hash_filter<K, CI, 64, RX>), dim3(grid), dim3(kFilterTPB), shmem, st, c, h, lit, cand);
  } else if (h.lane_bytes == 32) {
    hipError_t e = allow_lds(k_hash_filter<K, CI, 32, RX>, shmem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_hash_filter<K, CI, 32, RX>), dim3(grid), dim3(kFilterTPB), shmem, st, c, h, lit, cand);
  } else {
    hipError_t e = allow_lds(k_hash_filter<K, CI, 16, RX>, shmem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_hash_filter<K, CI, 16, RX>), dim3(grid), dim3(kFilterTPB), shmem, st, c, h, lit, cand);
  }
  return hipGetLastError();
}

template <int K, bool CI>
static hipError_t launch_hash_kc(uint32_t grid, hipStream_t st, const ScanCommon& c, const HashParams& h,
                                 const SliceOut& lit, const SliceOut& cand, bool rx) {
  return rx ? launch_hash_kcr<K, CI, true>(grid, st, c, h, lit, cand)
            : launch_hash_kcr<K, CI, false>(grid, st, c, h, lit, cand);
}

// The filter-and-verify kernel on `grid` blocks; block b writes output
// slice b of lit and cand (rx: the set has regex factors).
hipError_t launch_hash(uint32_t K, bool ci, uint32_t grid, hipStream_t st, const ScanCommon& c,
                       const HashParams& h, const SliceOut& lit, const SliceOut& cand, bool rx) {
  switch (K) {
#define WS_K(k)                                                                          \
  case k:                                                                                \
    return ci ? launch_hash_kc<k, true>(grid, st, c, h, lit, cand, rx)                  \
              : launch_hash_kc<k, false>(grid, st, c, h, lit, cand, rx);
    WS_K(1) WS_K(2) WS_K(3) WS_K(4) WS_K(6) WS_K(8) WS_K(12) WS_K(16)
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

hipError_t launch_chain_heads(const uint64_t* keys, const uint32_t* vals, const Rec* raw, uint32_t n,
                              const uint32_t* nd, uint32_t gshift, const uint32_t* maxlen, uint32_t maxlen_all,
                              uint8_t* head, hipStream_t st, const uint32_t* carry_src, uint32_t* carry_dst,
                              uint32_t ncarry) {
  hipLaunchKernelGGL(k_chain_heads, dim3(grid_of(n, nd)), dim3(kTPB), 0, st, keys, vals, raw, n, nd, gshift,
                     maxlen, maxlen_all, head, carry_src, carry_dst, ncarry);
  return hipGetLastError();
}

hipError_t launch_chain_walk(const uint64_t* keys, const uint32_t* vals, const Rec* raw, uint32_t n,
                             const uint32_t* nd, uint32_t gshift, const uint8_t* head, const uint32_t* carried,
                             uint32_t seg_off, const uint32_t* rx_pid, uint32_t ties, Rec* out,
                             uint32_t out_cap, uint32_t* out_count, uint32_t* carried_out,
                             hipStream_t st, const OrderOut* oo) {
  hipLaunchKernelGGL(k_chain_walk, dim3(grid_of(n, nd)), dim3(kTPB), 0, st, keys, vals, raw, n, nd, gshift, head,
                     carried, seg_off, rx_pid, ties, out, out_cap, out_count, carried_out, oo ? *oo : OrderOut{});
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
// the run's counters for the next run.
__global__ void k_run_close(uint32_t* counters, RunCaps caps, uint32_t* sticky, unsigned long long* total,
                            uint32_t* publish, uint32_t pwords) {
  if (threadIdx.x == 0) {
    uint32_t redo = 0;
    if (counters[5] > caps.lit_ovf_cap || counters[6] > caps.cand_ovf_cap) redo |= 2u;
    if (counters[kCounterLitSnapshot] > caps.lit_cap) redo |= 4u;
    if (counters[1] > caps.cand_cap) redo |= 8u;
    if (counters[2] > caps.raw_cap) redo |= 16u;
    if (counters[0] > caps.rec_cap) redo |= 32u;
    if (counters[9]) redo |= 64u;  // a sort bucket overflowed
    sticky[kStickyRedo] |= redo;
    if (counters[3] & kFlagTruncatedWalk) sticky[kStickyError] |= 1u;
    if (counters[0] > caps.max_matches) sticky[kStickyError] |= 2u;
    sticky[kStickyHits] += counters[8];
    sticky[kStickyCands] += counters[1];
    sticky[kStickyRaw] += counters[2];
    *total += counters[0];
    for (uint32_t k = 0; k <= kCounterLitSnapshot; ++k) counters[k] = 0;
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

__global__ void k_copy_u32(uint32_t* dst, const uint32_t* src) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *dst = *src;
}

hipError_t launch_run_close(uint32_t* counters, const RunCaps& caps, uint32_t* sticky,
                            unsigned long long* total, uint32_t* publish, uint32_t pwords, hipStream_t st) {
  hipLaunchKernelGGL(k_run_close, dim3(1), dim3(64), 0, st, counters, caps, sticky, total, publish, pwords);
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
__global__ __launch_bounds__(1024) void k_scan_counts(const uint32_t* cnt, uint32_t* cursor, uint32_t nb) {
  constexpr int T = kScanOneBlock / 4096;
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
    uint32_t x = own[t];
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, 64);
      if (lane >= (uint32_t)d) x += y;
    }
    inc[t] = x;
    if (lane == 63) s_tot[t * 16 + w] = x;
  }
  __syncthreads();
  if (w == 0) {
    const uint32_t a0 = s_tot[lane * 4], a1 = s_tot[lane * 4 + 1], a2 = s_tot[lane * 4 + 2], a3 = s_tot[lane * 4 + 3];
    const uint32_t sum = a0 + a1 + a2 + a3;
    uint32_t x = sum;
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, 64);
      if (lane >= (uint32_t)d) x += y;
    }
    const uint32_t ex = x - sum;
    s_tot[lane * 4] = ex;
    s_tot[lane * 4 + 1] = ex + a0;
    s_tot[lane * 4 + 2] = ex + a0 + a1;
    s_tot[lane * 4 + 3] = ex + a0 + a1 + a2;
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
      const uint32_t cb = __shfl(c, l, 64);
      const uint32_t bb = (uint32_t)__shfl((uint32_t)b, l, 64);
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
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, cnt, cursor, nb);
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
      const uint32_t lb = __shfl(b, leader);
      const uint64_t same = __ballot(live && b == lb) & todo;
      uint32_t base = 0;
      if ((int)lane_id() == leader) base = atomicAdd(&bcnt[lb], (uint32_t)__popcll(same));
      base = __shfl(base, leader);
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
  __shared__ uint32_t s_part[kGatherTPB];
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
  s_part[threadIdx.x] = sum;
  __syncthreads();
  for (uint32_t d = 1; d < kGatherTPB; d <<= 1) {
    const uint32_t x = threadIdx.x >= d ? s_part[threadIdx.x - d] : 0u;
    __syncthreads();
    s_part[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t run = s_part[threadIdx.x] - sum;
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
    counters[0] = s_part[threadIdx.x];
    counters[kCounterLitSnapshot] = s_part[threadIdx.x];
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
