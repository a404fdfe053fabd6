// The rolling-hash filter kernel (k_hash_filter) and the device helpers it
// shares with the other scan kernels, written for gfx950 (CDNA4, wave64).
// kernels.hip includes this header for the helpers; the filter's
// instantiations (one per window length K, case folding, lane width and
// regex flag) are compiled in filter_k*.hip, a few K per file, so they
// build in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "patterns.hpp"
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
  if constexpr (PL < 32) out[0] = mask >> (32 - PL);
}

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk2(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ u16x2 upk2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }

// K >= 3: the same filter over the 32-bit dual hash D (two 16-bit
// polynomial hashes rolled together by v_pk_mad_u16) of the window's K - 1
// little-endian byte pairs x[t] = b[t] | b[t+1] << 8. The pairs at even t
// are the halves of the loaded dwords, those at odd t the halves of one
// v_alignbyte per dword, and a multiply-add takes either half of a
// register (VOP3P operand selects). The bloom word's byte address is
// D & amask, its two bits D[20:16] and D[28:24], which the shifts take
// straight from the upper half and top byte of D (SDWA operand selects):
// per position two packed multiply-adds, a quarter of an align, one
// address op, the LDS read, two shifts, an and and the mask update.
// CONT: the window continues the lane's previous PL bytes, whose last
// window hash is D and whose last pair is xprev (the first position rolls
// on from them instead of hashing K - 1 pairs); on return D and xprev are
// this window's, for the next.
template <int K, bool CI, int PL, bool CONT>
__device__ __forceinline__ void filter_lane_dual(const HashParams& h, const uint32_t* bloom, uint32_t* d,
                                                 uint32_t* out, u16x2& D, u16x2& xprev) {
  constexpr int NW = PL / 4 + 4;
  static_assert(PL + K - 1 <= 4 * NW, "pair window bytes");
  if (CI) {
#pragma unroll
    for (int j = 0; j < NW; ++j) d[j] = fold_word(d[j]);
  }
  // od[m]: bytes 4m+1 .. 4m+4, the pairs at 4m+1 (low half) and 4m+3.
  uint32_t od[NW];
#pragma unroll
  for (int m = 0; m < NW; ++m) od[m] = __builtin_amdgcn_alignbyte(m + 1 < NW ? d[m + 1] : 0u, d[m], 1);
  auto xt = [&](int t) -> u16x2 {
    const u16x2 v = upk2((t & 1) ? od[t >> 2] : d[t >> 2]);
    return ((t >> 1) & 1) ? __builtin_shufflevector(v, v, 1, 1) : __builtin_shufflevector(v, v, 0, 0);
  };
  const u16x2 bb = upk2(h.BB), nbk = upk2(h.NBKK);
  uint32_t mask = 0;
  if constexpr (CONT) {
    D = D * bb + (xprev * nbk + xt(K - 2));
  } else {
    D = xt(0);
#pragma unroll
    for (int t = 1; t + 1 < K; ++t) D = D * bb + xt(t);
  }
  // The bloom words sit at LDS address 0 (k_hash_filter has no static LDS),
  // so the masked hash is the ds_read address as it stands.
  (void)bloom;
  const __attribute__((address_space(3))) char* lds = nullptr;
#pragma unroll
  for (int i = 0; i < PL; ++i) {
    if (i > 0) D = D * bb + (xt(i - 1) * nbk + xt(i + K - 2));
    const uint32_t Dw = pk2(D);
    const uint32_t w = *reinterpret_cast<const __attribute__((address_space(3))) uint32_t*>(lds + (Dw & h.amask));
    const uint32_t hit = (w >> ((Dw >> 16) & 31u)) & (w >> ((Dw >> 24) & 31u));
    mask = __builtin_amdgcn_alignbit(hit, mask, 1);
    if ((i & 31) == 31) out[i >> 5] = mask;
  }
  if constexpr (PL < 32) out[0] = mask >> (32 - PL);
  xprev = xt(PL - 1);
}

// Odd K (hash_stride(K) == 2): the windows at the even offsets p = 2i of
// the lane's PL bytes only (PL / 2 positions), rolled two bytes at a time:
// D(p + 2) = D(p) B^2 - B^K P(p) + P(p + K), P(t) = b[t] B + b[t + 1] (one
// packed multiply-add each, K odd so no P is used twice), then the same
// bloom test per window as above: half the bloom reads and tests per byte.
template <int K, bool CI, int PL>
__device__ __forceinline__ void filter_lane_dual2(const HashParams& h, uint32_t* d, uint32_t* out) {
  constexpr int NW = PL / 4 + 4;
  constexpr int NP = PL / 2;
  static_assert(K % 2 == 1 && K + PL - 2 < 4 * NW, "stride-2 window bytes");
  if (CI) {
#pragma unroll
    for (int j = 0; j < NW; ++j) d[j] = fold_word(d[j]);
  }
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
  const u16x2 bb = upk2(h.BB), bb2 = upk2(h.BB2), nbk = upk2(h.NBKK);
  auto P = [&](int t) -> u16x2 { return at(t) * bb + at(t + 1); };
  u16x2 D = at(0);
  uint32_t mask = 0;
#pragma unroll
  for (int t = 1; t < K; ++t) D = D * bb + at(t);
  const __attribute__((address_space(3))) char* lds = nullptr;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    if (i > 0) D = D * bb2 + (P(2 * i - 2) * nbk + P(2 * i - 2 + K));
    const uint32_t Dw = pk2(D);
    const uint32_t w = *reinterpret_cast<const __attribute__((address_space(3))) uint32_t*>(lds + (Dw & h.amask));
    const uint32_t hit = (w >> ((Dw >> 16) & 31u)) & (w >> ((Dw >> 24) & 31u));
    mask = __builtin_amdgcn_alignbit(hit, mask, 1);
    if ((i & 31) == 31) out[i >> 5] = mask;
  }
  if constexpr (NP < 32) out[0] = mask >> (32 - NP);
}

// Stride 2: valid window positions p0 + 2j (j < n <= 32): lo <= p,
// p + K <= hi, p - lo <= emit_end (the start p - 1 of a factor's second gram
// may still be below it; the verification checks each start).
__device__ __forceinline__ uint32_t valid_word2(uint32_t p0, uint32_t n, uint32_t lo, uint32_t hi, uint32_t K,
                                                uint32_t emit_end) {
  if (hi < K) return 0;
  const int64_t pmin = p0 > lo ? (int64_t)p0 : (int64_t)lo;
  int64_t pmax = (int64_t)hi - K;
  pmax = min(pmax, (int64_t)lo + (int64_t)emit_end);
  pmax = min(pmax, (int64_t)p0 + 2 * (int64_t)(n - 1));
  if (pmax < pmin) return 0;
  const uint32_t j0 = (uint32_t)((pmin - p0 + 1) >> 1), j1 = (uint32_t)((pmax - p0) >> 1) + 1;
  const uint32_t top = j1 >= 32 ? 0xFFFFFFFFu : ((1u << j1) - 1u);
  return top & ~((1u << j0) - 1u);
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

// v_ffbl_b32: the index of the lowest set bit, all ones for 0.
__device__ __forceinline__ uint32_t lowest_bit(uint32_t x) {
  uint32_t r;
  asm("v_ffbl_b32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// Exact comparison of one table entry (header {key, factor byte offset
// (16-byte aligned, gram offset cleared), length | kind << 31, output id},
// the factor's bytes 0-15 and 16-31) at position q. W holds the (folded,
// when CI) input bytes q..q+31. On a match *rec is the output record and
// the result is 1 (literal) or 2 (regex candidate); 0 otherwise.
// Every factor has at least K bytes: its first K are compared with masks
// known at compile time; in each further dword k of the entry the lowest
// differing bit must lie at or past the factor's end, 8 len - 32 k
// (clamped at 0): four operations a dword, no mask arithmetic. Factors
// longer than 32 bytes compare the rest from the factor bytes.
template <bool CI, int K>
__device__ __forceinline__ uint32_t verify_entry(const ScanCommon& c, uint4 e0, uint4 b0, uint4 b1,
                                                 const uint32_t* W, uint64_t q, uint2 seg, uint32_t item,
                                                 uint4* rec) {
  const uint32_t len = e0.z & 0x7FFFFFFFu;
  if (q + len > seg.y) return 0;
  const uint32_t pre[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
  uint32_t d0 = 0;
#pragma unroll
  for (int k = 0; 4 * k < K; ++k) {
    constexpr uint32_t kAll = 0xFFFFFFFFu;
    const uint32_t msk = 4 * k + 4 <= K ? kAll : ((1u << (8 * (K - 4 * k))) - 1u);
    d0 |= (W[k] ^ pre[k]) & msk;
  }
  if (d0) return 0;
  const uint32_t len8 = 8 * len;
  bool bad = false;
#pragma unroll
  for (int k = K / 4; k < 8; ++k) {
    const uint32_t lim = len8 > 32u * k ? len8 - 32u * k : 0u;
    bad |= lowest_bit(W[k] ^ pre[k]) < lim;
  }
  if (bad) return 0;
  const uint32_t start = (uint32_t)(q - seg.x);
  const uint32_t kind = (e0.z >> 31) ? 2u : 1u;
  // A literal match carries its end; a regex candidate its factor's slot
  // (byte offset / 16) + 1, where the walk can resume after the factor.
  *rec = make_uint4(e0.w, start, kind == 2u ? (e0.y >> 4) + 1u : start + len, item);
  if (len > 32) {
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
  base = __builtin_amdgcn_readlane(base, leader);
  if (take) slice_put(so, blockIdx.x, base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)), rec);
}

// A filter hit being verified: its position, item, segment, the input
// bytes from q (and, for stride-2 windows, from q - 1), and the window hash.
struct HitWork {
  uint64_t q;
  uint32_t item;
  uint2 seg;
  uint32_t W[8];
  uint32_t W1[8];
  uint32_t H;
  uint4 s0, s1, s2, s3;  // its bucket's entry: header, bytes 0-15, bytes 16-31, further entries
  bool live;             // false for a lane without a hit this round
};

// Bytes 3j .. 3j + 2 of the window in W (bytes at K and beyond zero) as a
// 24-bit value: one v_perm_b32.
template <int K>
__device__ __forceinline__ uint32_t key_chunk(const uint32_t* W, int j) {
  const int o = 3 * j, w0 = o >> 2;
  uint32_t sel = 0x0C000000u;  // byte 3: zero
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int b = o + i;
    // v_perm_b32 byte selectors: 0-3 the second operand (W[w0]), 4-7 the
    // first (W[w0 + 1]), 0x0C zero.
    const uint32_t si = b < K ? (uint32_t)(b - 4 * w0) : 0x0Cu;
    sel |= si << (8 * i);
  }
  const int last = o + 2 < K ? o + 2 : K - 1;  // the chunk's last window byte
  const bool hi = (last >> 2) > w0;
  return __builtin_amdgcn_perm(hi ? W[w0 + 1] : 0u, W[w0], sel);
}

// window_key (patterns.hpp) of the K bytes in W, already folded when CI.
template <int K>
__device__ __forceinline__ uint32_t window_key_dev(const uint32_t* W) {
  constexpr int NCH = (K + 2) / 3;
  uint32_t H = NCH == 1 ? key_chunk<K>(W, 0) : W[0];  // v_mad_u32_u24 reads the low 24 bits
#pragma unroll
  for (int j = 1; j < NCH; ++j) H = mad24_s(H, kWindowKeyMul, key_chunk<K>(W, j));
  return H & 0x7FFFFFFFu;
}

// raw: the nine dwords from q & ~3 (zeros past the loadable end), realigned
// with one v_alignbyte per dword.
// The window hash picks the bucket, whose two slots are read at once: one
// dependent read per hit. (Checking an 8-byte tag pair first and reading the
// slots only on a tag match was measured 1% slower on 10,000 DNA literals:
// the extra dependent read cost more than the slot bytes it saves.)
template <int K, bool CI>
__device__ __forceinline__ void hit_prepare(const ScanCommon& c, const HashParams& h, uint2 hr, uint2 seg,
                                            const uint32_t* raw, HitWork& hw) {
  hw.live = true;
  hw.q = hr.x;
  hw.item = hr.y;
  hw.seg = seg;
#pragma unroll
  for (int i = 0; i < 8; ++i) hw.W[i] = __builtin_amdgcn_alignbyte(raw[i + 1], raw[i], hr.x & 3u);
  if (CI) {
#pragma unroll
    for (int k = 0; k < 8; ++k) hw.W[k] = fold_word(hw.W[k]);
  }
  // The verification table's key of the window (its first 16 bytes).
  const uint32_t H = window_key_dev<K>(hw.W);
  hw.H = H;
  const uint4* slot = h.vtab + 4 * ((H * 0x9E3779B1u) >> (32 - h.nbits));
  hw.s0 = slot[0];
  hw.s1 = slot[1];
  hw.s2 = slot[2];
  hw.s3 = slot[3];
}

// Stride-2 windows: S = 2 hits are at window positions p, and an entry for
// a factor's gram at offset o (low bits of its byte-offset field) starts the
// factor at p - o, which must lie in the segment and below emit_end.
template <bool CI, int S, int K>
__device__ __forceinline__ uint32_t verify_gram(const ScanCommon& c, uint4 e0, uint4 b0, uint4 b1, const HitWork& hw,
                                                uint4* rec) {
  if constexpr (S == 1) {
    return verify_entry<CI, K>(c, e0, b0, b1, hw.W, hw.q, hw.seg, hw.item, rec);
  } else {
    const uint32_t o = e0.y & 15u;
    if (o > hw.q - hw.seg.x) return 0;
    const uint64_t q = hw.q - o;
    if (q - hw.seg.x >= c.emit_end) return 0;
    uint32_t W[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) W[i] = o ? hw.W1[i] : hw.W[i];
    e0.y &= ~15u;
    return verify_entry<CI, K>(c, e0, b0, b1, W, q, hw.seg, hw.item, rec);
  }
}

// RX: the set has regex factors (kind-2 entries, candidate output).
template <bool CI, bool RX, int K, int S = 1>
__device__ __forceinline__ void hit_finish(const ScanCommon& c, const HashParams& h, const HitWork& hw,
                                           uint32_t* s_n, const SliceOut& lit, const SliceOut& cand) {
  uint4 r0 = make_uint4(0, 0, 0, 0);
  uint32_t k0 = 0;
  if (hw.live && hw.s0.x == hw.H) k0 = verify_gram<CI, S, K>(c, hw.s0, hw.s1, hw.s2, hw, &r0);
  if (hw.live && hw.s3.y != 0) {
    // A bucket of two or more grams (rare): its entry lists the keys of the
    // first two further entries, so of those only the ones whose key is
    // the window's are read; entries past the second are all read.
    // Per-lane appends.
    uint32_t m = (uint32_t)(hw.s3.z == hw.H) | (uint32_t)(hw.s3.w == hw.H) << 1;
    const uint32_t n = hw.s3.y;
    // Visit order: the set bits of m, then indices 2 .. n - 1.
    for (uint32_t j = m ? __builtin_ctz(m) : 2u; j < n;) {
      const uint4* e = h.vmore + 4 * ((size_t)hw.s3.x + j);
      const uint4 e0 = e[0];
      if (e0.x == hw.H) {
        uint4 rec;
        const uint32_t k = verify_gram<CI, S, K>(c, e0, e[1], e[2], hw, &rec);
        if (k == 1) slice_put(lit, blockIdx.x, atomicAdd(&s_n[0], 1u), rec);
        if (RX && k == 2) slice_put(cand, blockIdx.x, atomicAdd(&s_n[1], 1u), rec);
      }
      m &= m - 1;
      j = m ? __builtin_ctz(m) : (j < 2 ? 2u : j + 1);
    }
  }
  emit_wave(lit, &s_n[0], k0 == 1, r0);
  if constexpr (RX) emit_wave(cand, &s_n[1], k0 == 2, r0);
}

// Per-wave LDS ring of filter hits waiting for verification (uint2
// {position, item}); a wave verifies 64 at a time, one per lane.
constexpr uint32_t kHitRing = 128;
// LDS byte offset of the hit rings: the bloom's end rounded up to 1 KiB
// (kHitRing slots of 8 bytes).
__host__ __device__ constexpr uint32_t hash_ring_offset(uint32_t nwords) {
  return (nwords * 4u + kHitRing * 8u - 1u) & ~(kHitRing * 8u - 1u);
}

// Verifies the n (<= 64) queued hits from ring index `head`, one per lane:
// the input dwords holding the window (streamed by the filter shortly
// before; nine single-dword loads from the window's dword, each naturally
// aligned, then one v_alignbyte per dword), the window key, its bucket's
// entry and the exact compare. (Three 16-byte loads of the 16-byte-aligned
// span need two bit-select steps more per dword; two unaligned 16-byte
// loads cost the memory pipeline more than the selects.)
// Call with the whole wave.
// seg0: the segment of a single-segment scan (c.units null), wave-uniform.
template <int K, bool CI, bool RX>
__device__ __forceinline__ void verify_queued(const ScanCommon& c, const HashParams& h, const uint2* ring,
                                              uint32_t head, uint32_t n, uint32_t* s_cnt, const SliceOut& lit,
                                              const SliceOut& cand, uint2 seg0, __amdgpu_buffer_rsrc_t rsrc) {
  const uint32_t lane = lane_id();
  HitWork hw;
  hw.live = false;
  hw.s0.x = kVtabEmpty;
  hw.s3.y = 0;
  if (lane < n) {
    const uint2 hr = ring[(head + lane) & (kHitRing - 1)];
    const uint2 seg = c.units ? c.items[hr.y] : seg0;
    const uint32_t a = hr.x & ~3u;
    uint32_t raw[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) raw[k] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, a + 4 * k, 0, 0);
    hit_prepare<K, CI>(c, h, hr, seg, raw, hw);
  }
  hit_finish<CI, RX, K>(c, h, hw, s_cnt, lit, cand);
}

// Stride-2 windows: the hit's span starts one byte early (at p - 1, or p at
// the buffer start) so that one load serves both candidate starts p and
// p - 1: bytes from p in W, from p - 1 in W1 (o1 = 1 when p > 0).
template <int K, bool CI>
__device__ __forceinline__ void hit_prepare_s2(const ScanCommon& c, const HashParams& h, uint2 hr, uint2 seg,
                                               const uint32_t* raw, uint32_t o1, HitWork& hw) {
  hw.live = true;
  hw.q = hr.x;
  hw.item = hr.y;
  hw.seg = seg;
  uint32_t Wb[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Wb[i] = __builtin_amdgcn_alignbyte(raw[i + 1], raw[i], (hr.x - o1) & 3u);
#pragma unroll
  for (int i = 0; i < 8; ++i) hw.W[i] = __builtin_amdgcn_alignbyte(Wb[i + 1], Wb[i], o1);
#pragma unroll
  for (int i = 0; i < 8; ++i) hw.W1[i] = Wb[i];
  if (CI) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      hw.W[k] = fold_word(hw.W[k]);
      hw.W1[k] = fold_word(hw.W1[k]);
    }
  }
  const uint32_t H = window_key_dev<K>(hw.W);
  hw.H = H;
  const uint4* slot = h.vtab + 4 * ((H * 0x9E3779B1u) >> (32 - h.nbits));
  hw.s0 = slot[0];
  hw.s1 = slot[1];
  hw.s2 = slot[2];
  hw.s3 = slot[3];
}

template <int K, bool CI, bool RX>
__device__ __forceinline__ void verify_queued_s2(const ScanCommon& c, const HashParams& h, const uint2* ring,
                                                 uint32_t head, uint32_t n, uint32_t* s_cnt, const SliceOut& lit,
                                                 const SliceOut& cand, uint2 seg0, __amdgpu_buffer_rsrc_t rsrc) {
  const uint32_t lane = lane_id();
  HitWork hw;
  hw.live = false;
  hw.s0.x = kVtabEmpty;
  hw.s3.y = 0;
  if (lane < n) {
    const uint2 hr = ring[(head + lane) & (kHitRing - 1)];
    const uint2 seg = c.units ? c.items[hr.y] : seg0;
    const uint32_t o1 = hr.x > 0 ? 1u : 0u;
    const uint32_t a = (hr.x - o1) & ~3u;
    uint32_t raw[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) raw[k] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, a + 4 * k, 0, 0);
    hit_prepare_s2<K, CI>(c, h, hr, seg, raw, o1, hw);
  }
  hit_finish<CI, RX, K, 2>(c, h, hw, s_cnt, lit, cand);
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

// Inclusive prefix sum over the wave: row_shr 1, 2, 4, 8 within each row of
// 16 lanes (DPP, lanes without a source add 0), then row_bcast:15 adds
// lane 15's sum to row 1 and lane 47's to row 3, and row_bcast:31 lane 31's
// to rows 2 and 3 (rows outside the row mask add 0): six DPP adds.
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t x, uint32_t lane) {
  (void)lane;
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);
  return x;
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
  // a word's address is the masked hash itself), one hit ring per wave from
  // the next 1 KiB boundary (a ring slot's address is the ring's base or'ed
  // with the slot's byte offset: one v_and_or_b32), then the block's two
  // output counters, its work counter and its hit count.
  extern __shared__ uint32_t smem[];
  uint32_t* bloom = smem;
  const uint32_t roff = hash_ring_offset(h.nwords);
  uint32_t* s_cnt = smem + (roff + kFilterTPB / 64 * kHitRing * 8) / 4;
  uint32_t* s_work = s_cnt + 2;
  uint32_t* s_hits = s_cnt + 3;
  const uint32_t ring_base = roff + (threadIdx.x / 64) * kHitRing * 8;  // LDS byte address
  uint2* ring = reinterpret_cast<uint2*>(reinterpret_cast<char*>(smem) + ring_base);
  __attribute__((address_space(3))) char* lds_w = nullptr;
  // Slot rb / 8 (mod kHitRing) of the wave's ring.
  typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
  const uint32_t slot_mask = kHitRing * 8 - 8;
  auto ring_put = [&](uint32_t rb, uint2 v) {
    // (The compiler turns the or of disjoint bits into an add, and then
    // spends two operations on it; v_and_or_b32 is one.)
    uint32_t a;
    asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(a) : "v"(rb), "s"(slot_mask), "v"(ring_base));
    u32x2 e;
    e.x = v.x;
    e.y = v.y;
    *reinterpret_cast<__attribute__((address_space(3))) u32x2*>(lds_w + a) = e;
  };
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
  // Half unit s of unit s / 2: 64 lanes x 2 PL bytes. At 16 positions per
  // lane (dual hash, every position) a lane's two quarters (j = 0, 1) are
  // contiguous and the second one's rolling hash goes on from the first's;
  // otherwise they are 64 PL bytes apart (64 lanes x PL bytes each).
  // Measured with the contiguous layout: the 100,000-item batch filter
  // (16 positions) 0.486 -> 0.440 ms, the 64 MiB regex chunks (32) 1% and
  // the 1 GiB k-mer filter (64) 12% slower.
  constexpr bool CONTIG = PL == 16 && K >= 3 && hash_stride(K) == 1;
  auto quarter_p0 = [&](const Half& hf, int j) {
    return CONTIG ? hf.tile_base + PL * (2 * (64 * (hf.s & 1) + lane) + j)
                  : hf.tile_base + PL * (64 * (2 * (hf.s & 1) + j) + lane);
  };
  // Granules outside the segment are clamped to its first or last one: one
  // load per granule on every path (so the waits before each use can count
  // them), and positions whose window leaves the segment are masked out
  // afterwards, whatever bytes they hashed.
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(c.base), 0, (int)load_end, 0x00020000);
  auto load_quarter = [&](const Half& hf, int j, uint32_t* dj) {
    const uint32_t p0 = quarter_p0(hf, j);
#pragma unroll
    for (int v = 0; v < NW / 4; ++v) {
      typedef int i4 __attribute__((ext_vector_type(4)));
      const i4 x = __builtin_amdgcn_raw_buffer_load_b128(rsrc, p0 + 16 * v, 0, 0);
      dj[4 * v] = (uint32_t)x.x; dj[4 * v + 1] = (uint32_t)x.y; dj[4 * v + 2] = (uint32_t)x.z; dj[4 * v + 3] = (uint32_t)x.w;
    }
  };
  // Window positions: every S-th byte of the lane's PL (S = 2 for odd K).
  constexpr int S = (int)hash_stride(K);
  constexpr int NPOS = PL / S;                    // windows per lane per quarter
  constexpr int NM = NPOS > 32 ? NPOS / 32 : 1;   // mask words
  constexpr uint32_t WPOS = NPOS < 32 ? NPOS : 32;  // windows per mask word
  // The dual hash's state between a lane's two contiguous quarters.
  u16x2 Dc = upk2(0u), xc = upk2(0u);
  auto filter_quarter = [&](const Half& hf, int j, uint32_t* dj, uint32_t* mj) {
    const uint32_t p0 = quarter_p0(hf, j);
#pragma unroll
    for (int k = 0; k < NM; ++k) mj[k] = 0;
    // Contiguous quarters: quarter 0 runs whenever either meets the segment
    // (quarter 1 rolls on from it).
    const uint32_t p1 = p0 + (CONTIG && j == 0 ? 2 * PL : PL);
    if (hf.inner || (p0 < hf.hi && p1 > hf.lo)) {
      if constexpr (S == 2)
        filter_lane_dual2<K, CI, PL>(h, dj, mj);
      else if constexpr (K >= 3) {
        if (CONTIG && j == 1)
          filter_lane_dual<K, CI, PL, CONTIG>(h, bloom, dj, mj, Dc, xc);
        else
          filter_lane_dual<K, CI, PL, false>(h, bloom, dj, mj, Dc, xc);
      } else
        filter_lane<K, CI, PL>(h, bloom, dj, mj);
      if (!hf.inner) {
#pragma unroll
        for (int k = 0; k < NM; ++k) {
          if constexpr (S == 2)
            mj[k] &= valid_word2(p0 + 2 * WPOS * k, WPOS, hf.lo, hf.hi, K, c.emit_end);
          else
            mj[k] &= valid_word(p0 + WPOS * k, WPOS, hf.lo, hf.hi, K, c.emit_end);
        }
      }
    }
  };
  auto verify = [&](uint32_t head, uint32_t n) {
    if constexpr (S == 2)
      verify_queued_s2<K, CI, RX>(c, h, ring, head, n, s_cnt, lit, cand, seg0, rsrc);
    else
      verify_queued<K, CI, RX>(c, h, ring, head, n, s_cnt, lit, cand, seg0, rsrc);
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
    uint32_t p0[kUnitsPerIter];
    const uint32_t item = cur.item;
    uint32_t left = 0;
#pragma unroll
    for (int j = 0; j < kUnitsPerIter; ++j) {
      p0[j] = quarter_p0(cur, j);
#pragma unroll
      for (int k = 0; k < NM; ++k) left |= mask[j][k];
    }
    cur_s = nxt_s;
    cur = nxt;
    if (more) {
      nxt_s = claim();
      nxt_rec = unit_rec(nxt_s);
    }
    // The hits to the wave's ring. Each lane's hits go at its rank among the
    // wave's T hits (a prefix sum of the lanes' counts) when the ring has
    // room for all T (pend < 64 here: always, for T <= 64). Otherwise a
    // round per hit: every lane with hits left puts its lowest one at its
    // rank among those lanes, the ring drained 64 at a time before a round
    // could overrun it.
    uint64_t live = __ballot(left != 0u);
    if (live) {
      uint32_t nh = 0;
#pragma unroll
      for (int j = 0; j < kUnitsPerIter; ++j)
#pragma unroll
        for (int k = 0; k < NM; ++k) nh += __builtin_popcount(mask[j][k]);
      const uint32_t x = wave_incl_sum(nh, lane);
      const uint32_t T = __builtin_amdgcn_readlane(x, 63);
      if (pend + T <= kHitRing) {
        uint32_t rb = (tail + x - nh) * 8u;  // the lane's first slot, in bytes
#pragma unroll
        for (int j = 0; j < kUnitsPerIter; ++j) {
#pragma unroll
          for (int k = 0; k < NM; ++k) {
            for (uint32_t m = mask[j][k]; m; m &= m - 1u, rb += 8u)
              ring_put(rb, make_uint2(p0[j] + S * (WPOS * k + __builtin_ctz(m)), item));
          }
        }
        tail += T;
        pend += T;
        nhits += T;
        live = 0;
      }
    }
    while (live) {
      if (pend >= 64) {
        wave_lds_sync();
        do {
          verify(tail - pend, 64);
          pend -= 64;
        } while (pend >= 64);
        wave_lds_sync();
      }
      if (left) {
        uint32_t pos = 0;
        bool found = false;
#pragma unroll
        for (int j = 0; j < kUnitsPerIter; ++j) {
#pragma unroll
          for (int k = 0; k < NM; ++k) {
            if (!found && mask[j][k]) {
              pos = p0[j] + S * (WPOS * k + __builtin_ctz(mask[j][k]));
              mask[j][k] &= mask[j][k] - 1;
              found = true;
            }
          }
        }
        const uint32_t r = __builtin_amdgcn_mbcnt_hi((uint32_t)(live >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)live, 0u));
        ring[(tail + r) & (kHitRing - 1)] = make_uint2(pos, item);
        left = 0;
#pragma unroll
        for (int j = 0; j < kUnitsPerIter; ++j)
#pragma unroll
          for (int k = 0; k < NM; ++k) left |= mask[j][k];
      }
      const uint32_t n = (uint32_t)__popcll(live);
      tail += n;
      pend += n;
      nhits += n;
      live = __ballot(left != 0u);
    }
    if (pend >= 64) {
      wave_lds_sync();
      do {
        verify(tail - pend, 64);
        pend -= 64;
      } while (pend >= 64);
      wave_lds_sync();
    }
  }
  wave_lds_sync();
  while (pend > 0) {
    const uint32_t n = min(pend, 64u);
    verify(tail - pend, n);
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


// Raises a kernel's dynamic LDS limit to `bytes` once (remembered per
// kernel), so launches inside a graph capture make no attribute calls.
hipError_t allow_lds_bytes(const void* fn, size_t bytes);
template <typename F>
static hipError_t allow_lds(F* fn, size_t bytes) {
  return allow_lds_bytes(reinterpret_cast<const void*>(fn), bytes);
}

inline size_t hash_shmem_bytes(uint32_t nwords) {
  return hash_ring_offset(nwords) + (kFilterTPB / 64) * kHitRing * sizeof(uint2) + 16u;
}

// Per window length K (filter_k*.hip): the LDS limit of its kernels, and a
// launch on `grid` blocks.
template <int K>
hipError_t prepare_hash_k(bool ci, uint32_t lane_bytes, size_t shmem, bool rx);
template <int K>
hipError_t launch_hash_k(bool ci, uint32_t grid, hipStream_t st, const ScanCommon& c, const HashParams& h,
                         const SliceOut& lit, const SliceOut& cand, bool rx);

#ifdef WS_FILTER_INSTANCES
template <int K, bool CI, bool RX>
static hipError_t prepare_hash_kcr(uint32_t lane_bytes, size_t shmem) {
  if (lane_bytes == 64) return allow_lds(k_hash_filter<K, CI, 64, RX>, shmem);
  if (lane_bytes == 32) return allow_lds(k_hash_filter<K, CI, 32, RX>, shmem);
  return allow_lds(k_hash_filter<K, CI, 16, RX>, shmem);
}

template <int K>
hipError_t prepare_hash_k(bool ci, uint32_t lane_bytes, size_t shmem, bool rx) {
  if (ci) return rx ? prepare_hash_kcr<K, true, true>(lane_bytes, shmem) : prepare_hash_kcr<K, true, false>(lane_bytes, shmem);
  return rx ? prepare_hash_kcr<K, false, true>(lane_bytes, shmem) : prepare_hash_kcr<K, false, false>(lane_bytes, shmem);
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

template <int K>
hipError_t launch_hash_k(bool ci, uint32_t grid, hipStream_t st, const ScanCommon& c, const HashParams& h,
                         const SliceOut& lit, const SliceOut& cand, bool rx) {
  if (ci) return rx ? launch_hash_kcr<K, true, true>(grid, st, c, h, lit, cand)
                    : launch_hash_kcr<K, true, false>(grid, st, c, h, lit, cand);
  return rx ? launch_hash_kcr<K, false, true>(grid, st, c, h, lit, cand)
            : launch_hash_kcr<K, false, false>(grid, st, c, h, lit, cand);
}

#define WS_FILTER_INSTANCE(k)                                                                                  \
  template hipError_t prepare_hash_k<k>(bool, uint32_t, size_t, bool);                                         \
  template hipError_t launch_hash_k<k>(bool, uint32_t, hipStream_t, const ScanCommon&, const HashParams&,      \
                                       const SliceOut&, const SliceOut&, bool);
#endif

}  // namespace ws
