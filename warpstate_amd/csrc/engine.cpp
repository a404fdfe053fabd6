// Host runtime and C ABI of the scan engine (GpuMatcher / StreamPipeline /
// scan_batch equivalents of the reference, core/src/gpu/mod.rs,
// core/src/pipeline.rs, core/src/batch.rs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <tuple>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/warpstate.h"
#include "index.hpp"
#include "kernels.hpp"
#include "patterns.hpp"

namespace ws {

// Launch wrappers defined in kernels.hip.
size_t hash_shmem_bytes(uint32_t nwords);
hipError_t launch_hash(uint32_t K, bool ci, uint32_t grid, hipStream_t st, const ScanCommon& c,
                       const HashParams& h, const SliceOut& lit, const SliceOut& cand, bool rx);
uint32_t ac_unit_bytes();
size_t ac_shmem_bytes(bool lds_table, uint32_t states, uint32_t classes);
hipError_t launch_ac(bool lds_table, uint32_t grid, hipStream_t st, const ScanCommon& c,
                     const AcParams& a);
hipError_t launch_walk(uint32_t ncand_host, hipStream_t st, const WalkParams& w);
hipError_t launch_walk_slices(hipStream_t st, const WalkParams& w, const SliceOut& cand, uint32_t nslices,
                              uint32_t* cand_total);
hipError_t launch_rx_scan(hipStream_t st, const WalkParams& w);
hipError_t launch_order_keys(const Rec* recs, uint32_t n, const uint32_t* nd, uint32_t sbits, uint32_t pbits,
                             uint64_t* keys, uint32_t* vals, hipStream_t st, const DenseOut& dense = DenseOut{});
hipError_t launch_chain_keys(const Rec* raw, uint32_t n, const uint32_t* nd, uint32_t sbits, uint32_t rbits,
                             uint64_t* keys, uint32_t* vals, hipStream_t st, const uint64_t* dbase = nullptr);
hipError_t launch_chain_walk(const uint64_t* keys, const uint32_t* vals, const Rec* raw, uint32_t n,
                             const uint32_t* nd, uint32_t gshift, const uint32_t* maxlen, uint32_t maxlen_all,
                             const uint32_t* carried, uint32_t* carried_out, uint32_t ncarry, uint32_t seg_off,
                             const uint32_t* rx_pid, uint32_t ties, Rec* out, uint32_t out_cap,
                             uint32_t* out_count, hipStream_t st, const OrderOut* oo = nullptr);
hipError_t launch_lit_keys(const Rec* recs, uint32_t n, uint32_t lbits, uint32_t pbits, uint32_t maxlen,
                           uint32_t len_desc, uint64_t* keys, uint32_t* vals, hipStream_t st);
hipError_t launch_pid_keys(const Rec* recs, uint32_t n, uint32_t sbits, uint64_t* keys, uint32_t* vals,
                           hipStream_t st);
hipError_t launch_pack_matches(const uint64_t* keys, const uint32_t* vals, uint32_t n, const uint32_t* nd,
                               uint32_t pbits, uint32_t start_off, void* out, uint64_t out_cap,
                               const unsigned long long* base_dev, hipStream_t st, const DenseOut& dense = DenseOut{});
hipError_t launch_pack_tagged(const uint64_t* keys, const uint32_t* vals, uint32_t n, uint32_t sbits,
                              uint32_t pbits, const uint64_t* ids, void* out, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t st);
hipError_t launch_slice_compact_keys(const SliceOut& lit, const SliceOut& cand, uint32_t n, uint32_t* lit_total,
                                     uint32_t* cand_total, uint32_t* lit_snap, uint32_t cap, uint32_t sbits,
                                     uint32_t pbits, uint64_t* keys, uint32_t* vals, uint32_t* cnt, uint32_t shift,
                                     Rec* cout, uint32_t ccap, hipStream_t st, const DenseOut& dense = DenseOut{});
hipError_t bucket_order_pack(void* tmp, size_t tmp_bytes, const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
                             uint32_t* vout, uint32_t n, const uint32_t* nd, uint32_t shift, uint32_t nb,
                             uint32_t* cnt, uint32_t* cursor, uint32_t* flag, uint32_t pbits, uint32_t start_off,
                             void* out, uint64_t out_cap, const unsigned long long* base_dev, hipStream_t st,
                             const DenseOut& dense = DenseOut{});
hipError_t launch_lit_order(const SliceOut& lit, uint32_t n, uint32_t pbits, uint32_t shift, uint32_t nb,
                            uint64_t* bkeys, uint32_t* bcnt, uint32_t* bofs, uint32_t* counters,
                            const uint32_t* flen, uint32_t start_off, void* out, uint64_t out_cap,
                            const unsigned long long* base_dev, hipStream_t st, const DenseOut& dense = DenseOut{});
hipError_t launch_order_keys_dense(const Rec* recs, uint32_t n, uint32_t pbits, const uint64_t* bases,
                                   uint64_t* keys, uint32_t* vals, hipStream_t st);
hipError_t launch_pack_tagged_dense(const uint64_t* keys, const uint32_t* vals, uint32_t n, uint32_t pbits,
                                    const uint64_t* bases, uint32_t nitems, const uint64_t* ids, void* out,
                                    hipStream_t st);
hipError_t launch_expand_units(const uint2* runs, uint32_t nruns, const uint2* items, uint32_t unit, uint4* units,
                               uint32_t nunits, hipStream_t st);
hipError_t launch_call_init(uint32_t* sticky, uint32_t nsticky, uint32_t* counters, uint32_t ncounters,
                            unsigned long long* stamps, uint32_t n, const CallItems& ci, uint2* items,
                            hipStream_t st);
hipError_t prepare_hash(uint32_t K, bool ci, uint32_t lane_bytes, uint32_t nwords, bool rx);
hipError_t prepare_ac(bool lds_table, uint32_t states, uint32_t classes);
hipError_t launch_lit_order_small(const SliceOut& lit, uint32_t n, uint32_t pbits, uint32_t sbits,
                                  const uint32_t* flen, uint32_t start_off, void* out, uint64_t out_cap,
                                  uint32_t* counters, uint32_t* fcnt, const RunCaps& caps, uint32_t* sticky,
                                  unsigned long long* total, uint32_t* publish, uint32_t pwords, hipStream_t st);
hipError_t launch_run_close(uint32_t* counters, uint32_t* fcnt, const RunCaps& caps, uint32_t* sticky,
                            unsigned long long* total, uint32_t* publish, uint32_t pwords, hipStream_t st);
hipError_t launch_copy_u32(uint32_t* dst, const uint32_t* src, hipStream_t st);
hipError_t launch_slice_compact2(const SliceOut& lit, const SliceOut& cand, uint32_t n, uint32_t* loffs,
                                 uint32_t* coffs, uint32_t* lit_total, uint32_t* cand_total, uint32_t* lit_snap,
                                 Rec* lout, uint32_t lcap, Rec* cout, uint32_t ccap, hipStream_t st);
hipError_t sort_pairs(void* tmp, size_t* tmp_bytes, const uint64_t* kin, uint64_t* kout,
                      const uint32_t* vin, uint32_t* vout, uint32_t n, uint32_t bits,
                      hipStream_t st);
hipError_t bucket_sort_pairs(void* tmp, size_t* tmp_bytes, const uint64_t* kin, uint64_t* kout,
                             const uint32_t* vin, uint32_t* vout, uint32_t n, const uint32_t* nd, uint32_t shift,
                             uint32_t nb, uint32_t* cnt, uint32_t* cursor, uint32_t* flag, hipStream_t st,
                             bool counted = false);

namespace {

thread_local std::string g_err;

ws_status fail(ws_status s, const std::string& msg) {
  g_err = msg;
  return s;
}

uint32_t bits_for(uint64_t v) { return v == 0 ? 1u : 64u - (uint32_t)__builtin_clzll(v); }

// Bumped whenever a workspace is reallocated: captured graphs hold raw
// device addresses, so a graph recorded under an older epoch is stale. Each
// context counts its own workspaces (DevBuf::epoch points at its counter),
// so one context's growth does not make another's graphs stale; buffers of
// no context (pattern blobs, batch plans) count here.
std::atomic<uint64_t> g_buf_epoch{1};
std::atomic<uint64_t> g_pattern_serial{1};

// Set while an allocation-free call runs on this thread: workspaces are then
// views into the static scratch array, and a request beyond a view's size is
// an error instead of a reallocation.
thread_local bool g_noalloc = false;

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  std::atomic<uint64_t>* epoch = &g_buf_epoch;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (g_noalloc) return hipErrorOutOfMemory;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    epoch->fetch_add(1);
    size_t want = std::max(bytes, (size_t)256);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Identity of a device scan's launch sequence: every kernel argument follows
// from these, so an equal key can replay a captured graph.
struct GraphKey {
  uint64_t pat_serial = 0, len = 0, chunk = 0, overlap = 0, cap = 0, epoch = 0;
  uint64_t batch = 0;  // batch plan serial (0: a single input or a stream)
  const void* data = nullptr;
  const void* out = nullptr;
  uint32_t out_mem = 0, variant = 0, lane = 0, bpc = 0, profiling = 0;
  uint32_t small = 0;  // literal-only runs ordered by the one-block kernel
  uint32_t wide = 0;   // sort buckets for dense runs (RunSetup ob / cb)
  bool operator<(const GraphKey& o) const {
    return std::tie(pat_serial, len, chunk, overlap, cap, epoch, batch, data, out, out_mem, variant, lane, bpc,
                    profiling, small, wide) < std::tie(o.pat_serial, o.len, o.chunk, o.overlap, o.cap, o.epoch,
                                                       o.batch, o.data, o.out, o.out_mem, o.variant, o.lane, o.bpc,
                                                       o.profiling, o.small, o.wide);
  }
};

struct GraphEntry {
  hipGraphExec_t exec = nullptr;
  uint32_t seen = 0;
  bool broken = false;  // capture failed once: stay eager
};

}  // namespace
}  // namespace ws

using namespace ws;

struct ws_ctx {
  int dev = 0;
  hipStream_t stream = nullptr;
  uint32_t cus = 0;
  ws_config cfg{};
  std::mutex mu;
  bool profiling = false;
  ws_timings last{};
  hipEvent_t ev[8] = {};
  hipEvent_t tev[2] = {};      // around a whole device-sized call (outside any graph)
  uint32_t wall_khz = 0;       // device wall-clock rate (kernel stamps)
  DevBuf recs, cands, raw, k0, v0, k1, v1, tmp, counters, items, units, input, outbuf,
      scratch, carried, ids, lslices, cslices, scounts, bcnt, bcur, dstate, k2, v2, bcnt2, bases, lkeys, lcnt, lofs,
      posb, lslices2, cslices2, scounts2;
  // Multi-run calls (stream chunks) on the hash variant filter run k + 1 on
  // stream2 while run k's later passes run on `stream`: its own slices
  // (lslices2 / cslices2 / scounts2 for odd runs) and counter block; xev:
  // fork, filter done (by parity), run done (by parity).
  hipStream_t stream2 = nullptr;
  hipEvent_t xev[5] = {};
  std::vector<hipEvent_t> pev;  // per-chunk profiling events of device-sized runs
  uint8_t* hstage = nullptr;    // pinned staging read by captured copies
  size_t hstage_cap = 0;
  size_t stage_used = 0;        // checked-pipeline calls: bytes of hstage in use
  uint32_t* hstate = nullptr;   // pinned, coherent: the device state block as the last kernel left it
  size_t hstate_cap = 0;
  // Device-sized calls: the state block, counters and segment list as the
  // last call left them (its close kernel resets the state block and the
  // counters; the segments are that call's). While `init_clean` holds, a
  // call with the same segments skips k_call_init. Every other user of the
  // counters or the segment list clears it.
  bool init_clean = false;
  // Literal-only device-sized runs order their matches with the one-block
  // kernel (k_lit_order_small) while the last such call had at most a
  // quarter of its capacity per run; a run with more is redone on the
  // checked path and turns it off.
  bool lit_small = false;
  bool graphs_on = true;  // capture repeated device-sized calls (ws_ctx_set_graphs)
  // Device-sized sorts use up to 2^14 buckets (a scan of their counts and
  // a rank grid cost a few us) unless the last call had more than 2^16
  // records per run or overflowed a bucket: then up to 2^16 (a bucket's
  // rank is quadratic in its size).
  bool sort_wide = false;
  CallItems clean_items{};
  const void* clean_state = nullptr;
  size_t clean_state_bytes = 0;
  const void* clean_items_buf = nullptr;
  std::map<GraphKey, GraphEntry> graphs;
  uint32_t* hcount = nullptr;  // pinned host mirror of the counters
  uint32_t last_hits = 0;      // filter hits of the last hash-variant scan
  // Workspace generation of this context (its DevBufs count here).
  std::atomic<uint64_t> epoch{1};
  // Stream chunks on the checked pipeline: which half of `carried` holds the
  // regex chains' state from the earlier chunks (the chain walk reads it and
  // raises the other half, then the halves swap).
  uint32_t carry_phase = 0;
  // The ws_scan_async call not yet collected by ws_wait, or null.
  struct ws_async_call* pending = nullptr;
};

struct ws_patterns {
  std::atomic<uint32_t> refs{1};
  int dev = 0;
  uint64_t serial = 0;  // unique per compiled set (keys captured graphs)
  PatternsHost h;
  DevBuf blob;
  // Device pointers into `blob`.
  const uint8_t* fbytes = nullptr;
  const uint4* fmeta = nullptr;
  const uint4 *vtab = nullptr, *vmore = nullptr;
  const uint32_t *foff = nullptr, *flen = nullptr, *fown = nullptr, *lit_pid = nullptr;
  const uint32_t* bloom = nullptr;
  const uint32_t *ac_table32 = nullptr, *ac_out_start = nullptr, *ac_out_list = nullptr,
                 *ac_len = nullptr;
  const uint16_t* ac_table16 = nullptr;
  const uint8_t* ac_cmap = nullptr;
  const uint32_t *rx_tables = nullptr, *rx_toff = nullptr, *rx_ncls = nullptr,
                 *rx_pid = nullptr, *rx_maxlen = nullptr, *rx_starts = nullptr, *rx_scan = nullptr;
  const uint8_t* rx_cmaps = nullptr;
  const uint2* rx_resume = nullptr;  // per 16-byte slot of fbytes x start context (regex factors; null: none)
  uint32_t rx_states_total = 0;
  bool ac_lds_ok = false;
};

namespace {

#define WS_HIP(expr)                                                                    \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess)                                                               \
      return fail(WS_ERR_GPU_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

constexpr size_t kAcLdsBudget = 96 * 1024;

// Builds one device blob holding every table of the compiled set.
ws_status upload_patterns(ws_patterns* p) {
  const PatternsHost& h = p->h;
  std::vector<uint8_t> blob;
  auto put = [&](const void* src, size_t bytes) -> size_t {
    size_t off = (blob.size() + 255) & ~(size_t)255;
    blob.resize(off + std::max<size_t>(bytes, 4));
    if (bytes) std::memcpy(blob.data() + off, src, bytes);
    return off;
  };
  std::string fb;
  std::vector<uint32_t> foff, flen, fown, aclen;
  for (const auto& f : h.fac) {
    while (fb.size() % 16) fb.push_back(0);  // 16-byte aligned factor bytes
    foff.push_back((uint32_t)fb.size());
    flen.push_back((uint32_t)f.bytes.size());
    fown.push_back((f.kind << 31) | f.owner);
    aclen.push_back(f.ac_len);
    for (char ch : f.bytes) {
      uint8_t b = (uint8_t)ch;
      if (h.ci && b >= 'A' && b <= 'Z') b += 32;
      fb.push_back((char)b);
    }
  }
  fb.append(32, '\0');  // 16-byte loads past the last factor stay in bounds
  std::vector<uint32_t> fmeta;
  for (size_t f = 0; f < h.fac.size(); ++f) {
    const auto& fc = h.fac[f];
    fmeta.push_back(foff[f]);
    fmeta.push_back(flen[f]);
    fmeta.push_back(fown[f]);
    fmeta.push_back(fc.kind ? fc.owner : h.lit_pid[fc.owner]);
  }
  // Verification table (see HashParams): one 64-byte entry per bucket
  // carrying the gram's key, factor byte offset | gram offset, length,
  // output id and first 32 (folded) bytes, so a filter hit is checked with
  // one 64-byte read; a bucket's further grams continue in `vmore`.
  std::vector<uint32_t> vtab, vmore;
  if (h.hash_ok) {
    const uint32_t nb = 1u << h.nbits;
    vtab.assign((size_t)nb * 16, 0);
    auto fill = [&](uint32_t* e, uint32_t ent) {
      const uint32_t f = h.ent_fac[ent];
      const Factor& fc = h.fac[f];
      e[0] = h.ent_hash[ent];
      e[1] = foff[f] | h.ent_off[ent];
      e[2] = flen[f] | (fc.kind << 31);
      e[3] = fc.kind ? fc.owner : h.lit_pid[fc.owner];
      uint8_t pre[32] = {0};
      std::memcpy(pre, fb.data() + foff[f], std::min<uint32_t>(32, flen[f]));
      std::memcpy(&e[4], pre, 32);
    };
    for (uint32_t b = 0; b < nb; ++b) {
      uint32_t* e = vtab.data() + (size_t)b * 16;
      e[0] = kVtabEmpty;
      const uint32_t e0 = h.bstart[b], n = h.bstart[b + 1] - e0;
      if (n >= 1) fill(e, h.bent[e0]);
      // {first, count, key 0, key 1} of the further entries in vmore
      // (kVtabEmpty: none; keys have their top bit clear).
      e[12] = (uint32_t)(vmore.size() / 16);
      e[13] = n > 1 ? n - 1 : 0;
      e[14] = n > 1 ? h.ent_hash[h.bent[e0 + 1]] : kVtabEmpty;
      e[15] = n > 2 ? h.ent_hash[h.bent[e0 + 2]] : kVtabEmpty;
      for (uint32_t k = 1; k < n; ++k) {
        vmore.resize(vmore.size() + 16);
        fill(vmore.data() + vmore.size() - 16, h.bent[e0 + k]);
      }
    }
  }
  size_t o_fmeta = put(fmeta.data(), fmeta.size() * 4);
  size_t o_vtab = put(vtab.data(), vtab.size() * 4);
  size_t o_vmore = put(vmore.data(), vmore.size() * 4);
  size_t o_fbytes = put(fb.data(), fb.size());
  size_t o_foff = put(foff.data(), foff.size() * 4);
  size_t o_flen = put(flen.data(), flen.size() * 4);
  size_t o_fown = put(fown.data(), fown.size() * 4);
  size_t o_aclen = put(aclen.data(), aclen.size() * 4);
  size_t o_litpid = put(h.lit_pid.data(), h.lit_pid.size() * 4);
  size_t o_bloom = put(h.bloom.data(), h.bloom.size() * 4);
  size_t o_ac32 = put(h.ac_table.data(), h.ac_table.size() * 4);
  std::vector<uint16_t> t16;
  p->ac_lds_ok = h.ac_ok && h.ac_states <= 65535 &&
                 (size_t)h.ac_states * h.ac_classes * 2 <= kAcLdsBudget;
  if (p->ac_lds_ok) {
    t16.resize(h.ac_table.size());
    for (size_t k = 0; k < t16.size(); ++k) t16[k] = (uint16_t)h.ac_table[k];
  }
  size_t o_ac16 = put(t16.data(), t16.size() * 2);
  size_t o_accmap = put(h.ac_cmap.data(), 256);
  size_t o_acos = put(h.ac_out_start.data(), h.ac_out_start.size() * 4);
  size_t o_acol = put(h.ac_out_list.data(), h.ac_out_list.size() * 4);
  std::vector<uint32_t> rxt, toff, ncls, rxmax, starts;
  std::vector<uint8_t> cm;
  p->rx_states_total = 0;
  for (size_t r = 0; r < h.rx.size(); ++r) {
    toff.push_back((uint32_t)rxt.size());
    ncls.push_back(h.rx[r].nclass + 1);  // row stride: the classes, then the end-of-text column
    starts.insert(starts.end(), h.rx[r].start.begin(), h.rx[r].start.end());
    rxt.insert(rxt.end(), h.rx[r].table.begin(), h.rx[r].table.end());
    cm.insert(cm.end(), h.rx[r].cmap.begin(), h.rx[r].cmap.end());
    long mx = h.rx_maxlen[r];
    rxmax.push_back(mx < 0 ? 0xFFFFFFFFu : (uint32_t)mx);
    p->rx_states_total += h.rx[r].nstates;
  }
  size_t o_rxt = put(rxt.data(), rxt.size() * 4);
  size_t o_toff = put(toff.data(), toff.size() * 4);
  size_t o_ncls = put(ncls.data(), ncls.size() * 4);
  size_t o_cm = put(cm.data(), cm.size());
  size_t o_rxpid = put(h.rx_pid.data(), h.rx_pid.size() * 4);
  size_t o_rxmax = put(rxmax.data(), rxmax.size() * 4);
  size_t o_rxstart = put(starts.data(), starts.size() * 4);
  size_t o_rxscan = put(h.rx_scan.data(), h.rx_scan.size() * 4);
  // Where an anchored walk stands once it has read a regex factor that the
  // prefilter verified at the walk's start: per factor (16-byte slot of its
  // bytes) and start context, {state (bit 31: dead inside the factor),
  // length << 16 | 1 + the offset of the last match flag inside it (0:
  // none)}. The walk then starts after the factor. Case-folding sets
  // verify folded bytes, so they walk from the start.
  std::vector<uint2> resume;
  if (!h.ci && !h.rx.empty()) {
    resume.assign((fb.size() / 16) * 4, make_uint2(0, 0));
    for (size_t f = 0; f < h.fac.size(); ++f) {
      const Factor& fc = h.fac[f];
      if (fc.kind != 1 || fc.bytes.size() > 0xFFFFu) continue;
      const AnchoredDfa& d = h.rx[fc.owner];
      const uint32_t C = d.nclass + 1;
      for (uint32_t cx = 0; cx < 4; ++cx) {
        uint32_t s = d.start[cx], last1 = 0;
        bool dead = false;
        for (size_t t = 0; t < fc.bytes.size(); ++t) {
          const uint32_t e = d.table[(size_t)s * C + d.cmap[(uint8_t)fc.bytes[t]]];
          if (e & DFA_MATCH) last1 = (uint32_t)t + 1;
          if (e & DFA_DEAD) {
            dead = true;
            break;
          }
          s = e & DFA_STATE_MASK;
        }
        resume[(foff[f] / 16) * 4 + cx] =
            make_uint2(dead ? 0x80000000u : s, ((uint32_t)fc.bytes.size() << 16) | last1);
      }
    }
  }
  size_t o_resume = put(resume.data(), resume.size() * sizeof(uint2));

  WS_HIP(p->blob.ensure(blob.size()));
  WS_HIP(hipMemcpy(p->blob.p, blob.data(), blob.size(), hipMemcpyHostToDevice));
  uint8_t* b = p->blob.as<uint8_t>();
  p->fbytes = b + o_fbytes;
  p->fmeta = reinterpret_cast<const uint4*>(b + o_fmeta);
  p->vtab = reinterpret_cast<const uint4*>(b + o_vtab);
  p->vmore = reinterpret_cast<const uint4*>(b + o_vmore);
  p->foff = reinterpret_cast<const uint32_t*>(b + o_foff);
  p->flen = reinterpret_cast<const uint32_t*>(b + o_flen);
  p->fown = reinterpret_cast<const uint32_t*>(b + o_fown);
  p->ac_len = reinterpret_cast<const uint32_t*>(b + o_aclen);
  p->lit_pid = reinterpret_cast<const uint32_t*>(b + o_litpid);
  p->bloom = reinterpret_cast<const uint32_t*>(b + o_bloom);
  p->ac_table32 = reinterpret_cast<const uint32_t*>(b + o_ac32);
  p->ac_table16 = reinterpret_cast<const uint16_t*>(b + o_ac16);
  p->ac_cmap = b + o_accmap;
  p->ac_out_start = reinterpret_cast<const uint32_t*>(b + o_acos);
  p->ac_out_list = reinterpret_cast<const uint32_t*>(b + o_acol);
  p->rx_tables = reinterpret_cast<const uint32_t*>(b + o_rxt);
  p->rx_toff = reinterpret_cast<const uint32_t*>(b + o_toff);
  p->rx_ncls = reinterpret_cast<const uint32_t*>(b + o_ncls);
  p->rx_cmaps = b + o_cm;
  p->rx_pid = reinterpret_cast<const uint32_t*>(b + o_rxpid);
  p->rx_maxlen = reinterpret_cast<const uint32_t*>(b + o_rxmax);
  p->rx_starts = reinterpret_cast<const uint32_t*>(b + o_rxstart);
  p->rx_scan = reinterpret_cast<const uint32_t*>(b + o_rxscan);
  p->rx_resume = resume.empty() ? nullptr : reinterpret_cast<const uint2*>(b + o_resume);
  return WS_OK;
}

uint32_t supported_variants(const ws_patterns* p) {
  uint32_t v = 1u << WS_VARIANT_HASH;
  if (p->h.fac.empty()) return v;  // no factors: no prefilter pass at all
  if (p->h.ac_ok) v |= 1u << WS_VARIANT_DFA_GLOBAL;
  if (p->ac_lds_ok) v |= 1u << WS_VARIANT_DFA_LDS;
  return v;
}

uint32_t pick_variant(const ws_patterns* p, const ws_tuning* t) {
  uint32_t v = t ? t->variant : WS_VARIANT_AUTO;
  if (v != WS_VARIANT_AUTO) return v;
  // Without a tuned choice: the hash prefilter unless the window is tiny and a
  // small factor DFA fits in LDS (a set without factors runs no prefilter).
  if (p->h.K <= 2 && p->ac_lds_ok && !p->h.fac.empty()) return WS_VARIANT_DFA_LDS;
  return WS_VARIANT_HASH;
}

// One scan over segments already laid out in device memory.
struct Plan {
  const uint8_t* base = nullptr;   // 16-aligned
  uint64_t load_end = 0;
  std::vector<uint2> items;        // {lo, hi}
  bool batch = false;              // explicit units per item
  uint32_t emit_end = 0xFFFFFFFFu;
  bool window_truncated = false;
  uint32_t seg_off = 0;            // absolute offset of item 0 (stream chunks)
  bool use_carried = false;
  bool ctx_before = false;         // base[items[0].x - 1] is the byte before the segment (a later stream chunk)
  uint64_t max_seg_len = 0;
  // A batch whose segments, work units (per lane width 16, 32, 64) and dense
  // positions are already in device memory (a batch plan): device-sized runs
  // take it as is.
  const uint2* dev_items = nullptr;
  const uint4* dev_units[3] = {};
  uint32_t dev_nunits[3] = {};
  DenseOut dense{};
  uint64_t total_bytes = 0;  // bytes of all items
  uint64_t total_pos = 0;    // dense positions: item bytes + one per item
};

uint32_t lane_index(uint32_t lane_bytes) { return lane_bytes == 64 ? 2u : lane_bytes == 32 ? 1u : 0u; }
// Which of a batch plan's unit lists a variant uses: the hash filter's for
// its lane width; the factor-DFA scan's 16 KiB units are the hash filter's at
// 64 positions per lane.
uint32_t unit_index(uint32_t variant, uint32_t lane_bytes) {
  return variant == WS_VARIANT_HASH ? lane_index(lane_bytes) : 2u;
}

// Candidate positions per regex of the every-position walk (k_rx_scan): each
// item's bytes and its end (an empty match may start there); a stream chunk
// takes the positions before its nominal end, and its end too when that is
// the text's end.
uint64_t scan_positions(const Plan& plan) {
  uint64_t n = 0;
  if (plan.batch) {
    for (const auto& it : plan.items) n += (uint64_t)(it.y - it.x) + 1;
    return n;
  }
  const uint64_t span = plan.items[0].y - plan.items[0].x;
  return plan.emit_end >= span ? span + 1 : plan.emit_end;
}

// The walk parameters of a plan (candidates, raw output and chain pairs are
// filled in by the caller).
WalkParams walk_params(const ws_patterns* p, const Plan& plan, const uint2* items, uint32_t* counters,
                       const uint64_t* pos_base) {
  WalkParams w{};
  w.base = plan.base;
  w.load_end = plan.load_end;
  w.items = items;
  w.tables = p->rx_tables;
  w.toff = p->rx_toff;
  w.ncls = p->rx_ncls;
  w.cmaps = p->rx_cmaps;
  w.resume = p->rx_resume;
  w.rx_pid = p->rx_pid;
  w.counters = counters;
  w.window_truncated = plan.window_truncated ? 1u : 0u;
  w.starts = p->rx_starts;
  w.ctx_before = plan.ctx_before ? 1u : 0u;
  w.scan_rx = p->rx_scan;
  w.nscan = (uint32_t)p->h.rx_scan.size();
  w.nitems = (uint32_t)plan.items.size();
  w.npos_total = w.nscan ? scan_positions(plan) : 0;
  w.pos_base = plan.batch ? pos_base : nullptr;
  return w;
}

struct Result {
  uint32_t count = 0;  // records in ctx->k1/v1 (sorted keys/values)
  uint32_t sbits = 0, pbits = 0;
  uint32_t keybits = 0;
  bool bucketed = false;  // ordered by the bucket sort: confirm after the next sync
  bool dense_batch = false;  // batch keys are (item base + start, pid): ctx->bases
};

ws_status sort_into(ws_ctx* ctx, uint32_t n, uint32_t bits, const uint64_t* kin, uint64_t* kout,
                    const uint32_t* vin, uint32_t* vout) {
  size_t need = 0;
  WS_HIP(sort_pairs(nullptr, &need, kin, kout, vin, vout, n, bits, ctx->stream));
  WS_HIP(ctx->tmp.ensure(need));
  size_t have = ctx->tmp.cap;
  WS_HIP(sort_pairs(ctx->tmp.p, &have, kin, kout, vin, vout, n, bits, ctx->stream));
  return WS_OK;
}

// Orders n (key, value) pairs from k0/v0 into k1/v1. Position-major match
// keys of a single input take the bucket sort (about two buckets per pair);
// its overflow flag (counters[9]) is copied to the host mirror with the next
// synchronization and checked by confirm_order.
// The bucket sort's counts: zero on allocation, and every sort leaves them
// zero again.
hipError_t ensure_zeroed_counts(ws_ctx* ctx, DevBuf& b, uint32_t nb) {
  const size_t bytes = (size_t)nb * 4;
  if (bytes <= b.cap) return hipSuccess;
  hipError_t e = b.ensure(bytes);
  if (e != hipSuccess) return e;
  return hipMemsetAsync(b.p, 0, b.cap, ctx->stream);
}
// A bucket sort's counts: at least kScanOneBlock of them (the one-block
// count scan reads that many).
hipError_t ensure_bucket_counts(ws_ctx* ctx, DevBuf& b, uint32_t nb) {
  return ensure_zeroed_counts(ctx, b, std::max(nb, kScanOneBlock));
}
hipError_t ensure_bucket_counts(ws_ctx* ctx, uint32_t nb) { return ensure_bucket_counts(ctx, ctx->bcnt, nb); }

ws_status order_pairs(ws_ctx* ctx, uint32_t n, uint32_t bits, bool dense, Result* res) {
  res->bucketed = false;
  res->keybits = bits;
  if (!dense || n < 4096 || bits > 63)
    return sort_into(ctx, n, bits, ctx->k0.as<uint64_t>(), ctx->k1.as<uint64_t>(), ctx->v0.as<uint32_t>(),
                     ctx->v1.as<uint32_t>());
  const uint32_t bb = std::min<uint32_t>(bits, std::min(22u, std::max(10u, bits_for(n) + 1)));
  const uint32_t shift = bits - bb, nb = 1u << bb;
  WS_HIP(ensure_bucket_counts(ctx, nb));
  WS_HIP(ctx->bcur.ensure((size_t)std::max(nb, kScanOneBlock) * 4));
  size_t need = 0;
  WS_HIP(bucket_sort_pairs(nullptr, &need, nullptr, nullptr, nullptr, nullptr, n, nullptr, shift, nb,
                           ctx->bcnt.as<uint32_t>(), ctx->bcur.as<uint32_t>(), nullptr, ctx->stream));
  WS_HIP(ctx->tmp.ensure(need));
  size_t have = ctx->tmp.cap;
  WS_HIP(launch_fill_u32(ctx->counters.as<uint32_t>() + 9, 1, 0, ctx->stream));
  WS_HIP(bucket_sort_pairs(ctx->tmp.p, &have, ctx->k0.as<uint64_t>(), ctx->k1.as<uint64_t>(),
                           ctx->v0.as<uint32_t>(), ctx->v1.as<uint32_t>(), n, nullptr, shift, nb,
                           ctx->bcnt.as<uint32_t>(), ctx->bcur.as<uint32_t>(),
                           ctx->counters.as<uint32_t>() + 9, ctx->stream));
  WS_HIP(hipMemcpyAsync(ctx->hcount + 9, ctx->counters.as<uint32_t>() + 9, 4, hipMemcpyDeviceToHost,
                        ctx->stream));
  res->bucketed = true;
  return WS_OK;
}

// After a synchronization that follows order_pairs: if a bucket overflowed,
// redo the ordering with the radix sort. Returns whether it did.
ws_status confirm_order(ws_ctx* ctx, Result& r, bool* redone) {
  ctx->init_clean = false;
  *redone = false;
  if (!r.bucketed) return WS_OK;
  r.bucketed = false;
  if (!ctx->hcount[9]) return WS_OK;
  ws_status s = sort_into(ctx, r.count, r.keybits, ctx->k0.as<uint64_t>(), ctx->k1.as<uint64_t>(),
                          ctx->v0.as<uint32_t>(), ctx->v1.as<uint32_t>());
  if (s != WS_OK) return s;
  *redone = true;
  return WS_OK;
}

ws_status read_counters(ws_ctx* ctx) {
  ctx->init_clean = false;
  WS_HIP(hipMemcpyAsync(ctx->hcount, ctx->counters.p, 9 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                        ctx->stream));
  WS_HIP(hipStreamSynchronize(ctx->stream));
  return WS_OK;
}

// One greedy pass over records sorted by `keys/vals` (records read from
// ctx->scratch, kept ones appended to ctx->recs from index 0).
ws_status greedy_pass(ws_ctx* ctx, uint32_t n, uint32_t gshift, uint32_t maxlen, uint32_t ties,
                      uint32_t* kept) {
  ctx->init_clean = false;
  const hipStream_t st = ctx->stream;
  WS_HIP(launch_fill_u32(ctx->counters.as<uint32_t>(), 1, 0, st));
  WS_HIP(launch_chain_walk(ctx->k1.as<uint64_t>(), ctx->v1.as<uint32_t>(), ctx->scratch.as<Rec>(), n, nullptr,
                           gshift, nullptr, maxlen, nullptr, nullptr, 0, 0, nullptr, ties,
                           ctx->recs.as<Rec>(), n, ctx->counters.as<uint32_t>(), st));
  ws_status s = read_counters(ctx);
  if (s != WS_OK) return s;
  *kept = ctx->hcount[0];
  return WS_OK;
}

// Non-overlapping literal semantics of PatternSet::scan, derived exactly from
// the overlapping match list in ctx->recs[0, nlit):
//  * LeftmostLongest (Aho-Corasick find_iter, core/src/cpu/scan.rs:250-293):
//    order by (start, longest, lowest id), keep greedily.
//  * MultiMemchr (core/src/specialize.rs:313-373): each needle's own
//    non-overlapping occurrences, then the same greedy order.
//  * HashScanner (core/src/hash_scan.rs:321-363): at the cursor every id of the
//    shortest matching length, then jump to its end.
ws_status reduce_literals(ws_ctx* ctx, uint32_t strategy, uint32_t nlit, uint32_t sbits,
                          uint32_t pbits, uint32_t maxlit) {
  const hipStream_t st = ctx->stream;
  const uint32_t lbits = bits_for(maxlit);
  if (sbits + lbits + pbits > 64)
    return fail(WS_ERR_INPUT_TOO_LARGE, "reduction key does not fit 64 bits");
  uint32_t n = nlit;
  if (strategy == STRAT_MULTI_MEMCHR) {
    WS_HIP(hipMemcpyAsync(ctx->scratch.p, ctx->recs.p, (size_t)n * sizeof(Rec), hipMemcpyDeviceToDevice, st));
    WS_HIP(launch_pid_keys(ctx->scratch.as<Rec>(), n, sbits, ctx->k0.as<uint64_t>(), ctx->v0.as<uint32_t>(), st));
    ws_status s = sort_into(ctx, n, sbits + pbits, ctx->k0.as<uint64_t>(), ctx->k1.as<uint64_t>(),
                            ctx->v0.as<uint32_t>(), ctx->v1.as<uint32_t>());
    if (s != WS_OK) return s;
    s = greedy_pass(ctx, n, sbits, maxlit, 0, &n);
    if (s != WS_OK) return s;
  }
  WS_HIP(hipMemcpyAsync(ctx->scratch.p, ctx->recs.p, (size_t)n * sizeof(Rec), hipMemcpyDeviceToDevice, st));
  const uint32_t desc = strategy == STRAT_HASH_SHORTEST ? 0u : 1u;
  WS_HIP(launch_lit_keys(ctx->scratch.as<Rec>(), n, lbits, pbits, maxlit, desc, ctx->k0.as<uint64_t>(),
                         ctx->v0.as<uint32_t>(), st));
  ws_status s = sort_into(ctx, n, sbits + lbits + pbits, ctx->k0.as<uint64_t>(), ctx->k1.as<uint64_t>(),
                          ctx->v0.as<uint32_t>(), ctx->v1.as<uint32_t>());
  if (s != WS_OK) return s;
  s = greedy_pass(ctx, n, 64, maxlit, strategy == STRAT_HASH_SHORTEST ? 1u : 0u, &n);
  if (s != WS_OK) return s;
  ctx->hcount[0] = n;
  return WS_OK;
}

// Hash filter: 16, 32 or 64 positions per lane. More positions amortise the
// per-window setup (K-1 hash steps, loads, bounds) but make work units
// coarser: 64 for large inputs, 16 for batches of small items.
uint32_t choose_lane_bytes(const Plan& plan, const ws_tuning* tuning) {
  uint32_t lane_bytes = tuning && tuning->bytes_per_lane ? tuning->bytes_per_lane : 0;
  if (lane_bytes == 16 || lane_bytes == 32 || lane_bytes == 64) return lane_bytes;
  uint64_t span = 0;
  for (auto& it : plan.items) span += it.y - it.x;
  if (plan.batch && span / plan.items.size() < 64 * 1024) return 16;
  // Measured on 16 MiB (16: 13.3 us, 32: 14.2, 64: 17.1), 64 MiB (32 = 64)
  // and 256 MiB (64: 88 us, 32: 98) single inputs.
  if (span <= (16ull << 20)) return 16;
  return span >= (256ull << 20) ? 64 : 32;
}

// Work units of `unit` bytes. A single input strides statically; a batch
// lists its units in work order: largest items first (longest-processing-
// time first, by power-of-two size class: a counting sort, O(items)), each
// item's tiles contiguous. The host keeps one run per item, {item, first
// unit}; k_expand_units writes the unit records on the device.
uint32_t plan_units(const Plan& plan, uint32_t unit, std::vector<uint2>* runs) {
  if (!plan.batch) return (uint32_t)((plan.items[0].y + unit - 1) / unit);
  const size_t ni = plan.items.size();
  auto cls = [&](size_t i) {
    const uint32_t len = plan.items[i].y - plan.items[i].x;
    return 32u - (uint32_t)__builtin_clz(len | 1u);  // 1..32
  };
  std::vector<uint32_t> start(34, 0);
  for (size_t i = 0; i < ni; ++i) start[33 - cls(i)]++;
  for (uint32_t k = 0, acc = 0; k < 34; ++k) { const uint32_t c = start[k]; start[k] = acc; acc += c; }
  std::vector<uint32_t> by_size(ni);
  for (size_t i = 0; i < ni; ++i) by_size[start[33 - cls(i)]++] = (uint32_t)i;
  runs->reserve(ni);
  uint32_t first = 0;
  for (uint32_t i : by_size) {
    const uint64_t lo = plan.items[i].x & ~15u, hi = plan.items[i].y;
    if (hi <= lo) continue;
    runs->push_back(make_uint2(i, first));
    first += (uint32_t)((hi - lo + unit - 1) / unit);
  }
  return first;
}

// Per-block output slices over `buf` (`total` records): a quarter goes to the
// shared overflow area.
SliceOut make_slices(DevBuf& buf, uint64_t total, uint32_t nvb, uint32_t* counts, uint32_t* ovf_count) {
  SliceOut so{};
  so.ovf_cap = (uint32_t)std::min<uint64_t>(total / 4, 0xFFFFFFF0ull);
  so.cap = (uint32_t)((total - so.ovf_cap) / nvb);
  so.buf = buf.as<Rec>();
  so.ovf = so.buf + (size_t)so.cap * nvb;
  so.counts = counts;
  so.ovf_count = ovf_count;
  return so;
}

// Records per output list (literal matches or regex candidates) of the
// hash variant: 1024 per filter block plus a shared overflow area of 256 K.
uint64_t slice_records(uint32_t nblocks) { return (uint64_t)nblocks * 1024 + (1u << 18); }

ScanCommon make_common(ws_ctx* ctx, const ws_patterns* p, const Plan& plan, uint32_t nunits, uint32_t lit_cap,
                       uint32_t cand_cap) {
  ScanCommon c{};
  c.base = plan.base;
  c.load_end = plan.load_end;
  c.units = plan.batch ? ctx->units.as<uint4>() : nullptr;
  c.nunits = nunits;
  c.items = ctx->items.as<uint2>();
  c.emit_end = plan.emit_end;
  c.ci = p->h.ci ? 1u : 0u;
  c.fbytes = p->fbytes;
  c.foff = p->foff;
  c.flen = p->flen;
  c.fown = p->fown;
  c.lit_pid = p->lit_pid;
  c.lit_out = ctx->recs.as<Rec>();
  c.lit_cap = lit_cap;
  c.cand_out = ctx->cands.as<Rec>();
  c.cand_cap = cand_cap;
  c.counters = ctx->counters.as<uint32_t>();
  // Batches pull uneven work from a device queue (zeroed with the counters);
  // a single input's equal tiles use a static stride.
  c.queue = plan.batch ? ctx->counters.as<uint32_t>() + 7 : nullptr;
  return c;
}

HashParams hash_params(const ws_patterns* p, uint32_t lane_bytes) {
  const PatternsHost& h = p->h;
  HashParams hp{};
  hp.bloom = p->bloom;
  hp.nwords = (uint32_t)h.bloom.size();
  hp.wshift = h.wshift;
  hp.wbits = h.wbits;
  hp.B = h.B;
  hp.NBK = h.NBK;
  hp.nbits = h.nbits;
  hp.s2 = h.s2;
  hp.lane_bytes = lane_bytes;
  hp.BB = h.BB;
  hp.BB2 = h.BB2;
  hp.NBKK = h.NBKK;
  hp.amask = (hp.nwords - 1u) << 2;
  hp.vtab = p->vtab;
  hp.vmore = p->vmore;
  hp.fmeta = p->fmeta;
  return hp;
}

hipError_t ensure_hstate(ws_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->hstate_cap) return hipSuccess;
  if (ctx->hstate) (void)hipHostFree(ctx->hstate);
  ctx->hstate = nullptr;
  ctx->hstate_cap = 0;
  ctx->epoch.fetch_add(1);
  const size_t want = std::max<size_t>(bytes, 4096);
  hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&ctx->hstate), want, hipHostMallocCoherent);
  if (e == hipSuccess) ctx->hstate_cap = want;
  return e;
}

hipError_t ensure_stage(ws_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->hstage_cap) return hipSuccess;
  // Copies from the old buffer may still be in flight on the stream.
  if (ctx->hstage) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return e;
    (void)hipHostFree(ctx->hstage);
  }
  ctx->hstage = nullptr;
  ctx->hstage_cap = 0;
  ctx->epoch.fetch_add(1);
  const size_t want = std::max<size_t>(bytes, 4096);
  hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&ctx->hstage), want, hipHostMallocDefault);
  if (e == hipSuccess) ctx->hstage_cap = want;
  return e;
}

// Host-to-device copy through the pinned staging area (pageable copies are
// slower and vary). ctx->stage_used is the running offset of the call's
// staged copies (reset when a checked-pipeline call starts).
hipError_t h2d_staged(ws_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!bytes) return hipSuccess;
  const size_t off = (ctx->stage_used + 255) & ~(size_t)255;
  hipError_t e = ensure_stage(ctx, off + bytes);
  if (e != hipSuccess) return e;
  std::memcpy(ctx->hstage + off, src, bytes);
  ctx->stage_used = off + bytes;
  return hipMemcpyAsync(dst, ctx->hstage + off, bytes, hipMemcpyHostToDevice, ctx->stream);
}

ws_status run_plan(ws_ctx* ctx, const ws_patterns* p, const Plan& plan, uint32_t mode,
                   const ws_tuning* tuning, Result* res) {
  ctx->init_clean = false;
  const PatternsHost& h = p->h;
  const hipStream_t st = ctx->stream;
  const uint32_t variant = pick_variant(p, tuning);
  if (!(supported_variants(p) & (1u << variant)))
    return fail(WS_ERR_INVALID_ARGUMENT, "kernel variant not available for this pattern set");
  const uint32_t lane_bytes = choose_lane_bytes(plan, tuning);
  const uint32_t unit = variant == WS_VARIANT_HASH ? kTPB * lane_bytes : ac_unit_bytes();
  std::vector<uint2> runs;
  const uint32_t nunits = plan_units(plan, unit, &runs);
  ctx->stage_used = 0;
  WS_HIP(ctx->items.ensure(plan.items.size() * sizeof(uint2)));
  WS_HIP(h2d_staged(ctx, ctx->items.p, plan.items.data(), plan.items.size() * sizeof(uint2)));
  if (plan.batch && nunits) {
    WS_HIP(ctx->units.ensure((size_t)nunits * sizeof(uint4) + runs.size() * sizeof(uint2)));
    uint2* druns = reinterpret_cast<uint2*>(ctx->units.as<uint4>() + nunits);
    WS_HIP(h2d_staged(ctx, druns, runs.data(), runs.size() * sizeof(uint2)));
    WS_HIP(launch_expand_units(druns, (uint32_t)runs.size(), ctx->items.as<uint2>(), unit, ctx->units.as<uint4>(),
                               nunits, st));
  }
  const uint32_t max_matches = (uint32_t)std::min<uint64_t>(ctx->cfg.max_matches, 0xFFFFFFF0ull);
  uint32_t lit_cap = (uint32_t)std::min<uint64_t>(ctx->recs.cap / sizeof(Rec), max_matches);
  uint32_t cand_cap = (uint32_t)(ctx->cands.cap / sizeof(Rec));
  if (lit_cap < 4096) {
    WS_HIP(ctx->recs.ensure(std::min<uint64_t>(max_matches, 1u << 16) * sizeof(Rec)));
    lit_cap = (uint32_t)std::min<uint64_t>(ctx->recs.cap / sizeof(Rec), max_matches);
  }
  if (cand_cap < 4096) {
    WS_HIP(ctx->cands.ensure((1u << 16) * sizeof(Rec)));
    cand_cap = (uint32_t)(ctx->cands.cap / sizeof(Rec));
  }
  const uint32_t per_cu = tuning && tuning->blocks_per_cu ? tuning->blocks_per_cu : 4;
  const uint32_t grid = std::max(1u, std::min(nunits, ctx->cus * per_cu));
  // Hash variant: one 1024-lane filter block per CU (its LDS holds the bloom
  // filter); each block owns a slice of the literal and candidate outputs,
  // plus a shared overflow area.
  const uint32_t nvb = std::max(1u, std::min((nunits + 7) / 8, ctx->cus * (per_cu > 4 ? 2u : 1u)));
  uint64_t lslice_total = ctx->lslices.cap / sizeof(Rec), cslice_total = ctx->cslices.cap / sizeof(Rec);
  if (variant == WS_VARIANT_HASH) {
    const uint64_t want = slice_records(nvb);
    if (lslice_total < want) { WS_HIP(ctx->lslices.ensure(want * sizeof(Rec))); lslice_total = want; }
    if (cslice_total < want) { WS_HIP(ctx->cslices.ensure(want * sizeof(Rec))); cslice_total = want; }
    WS_HIP(ctx->scounts.ensure(((size_t)nvb * 2 + 4) * 4 * 2));
  }
  uint32_t lit_ovf_cap = 0, cand_ovf_cap = 0;
  if (ctx->profiling) WS_HIP(hipEventRecord(ctx->ev[0], st));
  for (int attempt = 0;; ++attempt) {
    WS_HIP(launch_fill_u32(ctx->counters.as<uint32_t>(), 9, 0, st));
    ScanCommon c = make_common(ctx, p, plan, nunits, lit_cap, cand_cap);
    if (ctx->profiling) WS_HIP(hipEventRecord(ctx->ev[1], st));
    // A set whose regexes all scan every position has no factors: no filter.
    const bool filter = nunits > 0 && !h.fac.empty();
    if (!filter && ctx->profiling) WS_HIP(hipEventRecord(ctx->ev[6], st));
    if (filter) {
      if (variant == WS_VARIANT_HASH) {
        const HashParams hp = hash_params(p, lane_bytes);
        uint32_t* sc = ctx->scounts.as<uint32_t>();
        SliceOut lso = make_slices(ctx->lslices, lslice_total, nvb, sc, ctx->counters.as<uint32_t>() + 5);
        SliceOut cso = make_slices(ctx->cslices, cslice_total, nvb, sc + nvb, ctx->counters.as<uint32_t>() + 6);
        WS_HIP(launch_hash(h.K, h.ci, nvb, st, c, hp, lso, cso, !h.rx.empty()));
        if (ctx->profiling) WS_HIP(hipEventRecord(ctx->ev[6], st));
        uint32_t* offs = sc + 2 * nvb;
        WS_HIP(launch_slice_compact2(lso, cso, nvb, offs, offs + nvb + 2, ctx->counters.as<uint32_t>() + 0,
                                     ctx->counters.as<uint32_t>() + 1, nullptr, c.lit_out, c.lit_cap, c.cand_out,
                                     c.cand_cap, st));
        lit_ovf_cap = lso.ovf_cap;
        cand_ovf_cap = cso.ovf_cap;
      } else {
        AcParams ap{p->ac_table32, p->ac_table16, p->ac_cmap, p->ac_out_start, p->ac_out_list,
                    p->ac_len, h.ac_states, h.ac_classes, h.ac_first_out, h.ac_maxlen};
        WS_HIP(launch_ac(variant == WS_VARIANT_DFA_LDS, grid, st, c, ap));
        if (ctx->profiling) WS_HIP(hipEventRecord(ctx->ev[6], st));
      }
    }
    if (ctx->profiling) WS_HIP(hipEventRecord(ctx->ev[2], st));
    ws_status s = read_counters(ctx);
    if (s != WS_OK) return s;
    const uint32_t nlit = ctx->hcount[0], ncand = ctx->hcount[1];
    ctx->last_hits = 0;
    if (nlit > max_matches) {
      res->count = nlit;
      return fail(WS_ERR_MATCH_BUFFER_OVERFLOW,
                  "too many matches (" + std::to_string(nlit) + " exceeds " +
                      std::to_string(max_matches) + ")");
    }
    // Grow every buffer that overflowed, then rescan; each category grows
    // past its need at once, so this converges in one or two rounds.
    bool retry = false;
    if (variant == WS_VARIANT_HASH && ctx->hcount[5] > lit_ovf_cap) {
      lslice_total = lslice_total * 2 + (uint64_t)ctx->hcount[5] * 8;
      WS_HIP(ctx->lslices.ensure(lslice_total * sizeof(Rec)));
      retry = true;
    }
    if (variant == WS_VARIANT_HASH && ctx->hcount[6] > cand_ovf_cap) {
      cslice_total = cslice_total * 2 + (uint64_t)ctx->hcount[6] * 8;
      WS_HIP(ctx->cslices.ensure(cslice_total * sizeof(Rec)));
      retry = true;
    }
    if (nlit > lit_cap) {
      WS_HIP(ctx->recs.ensure((size_t)nlit * 2 * sizeof(Rec) > (size_t)max_matches * sizeof(Rec) * 2
                                  ? (size_t)max_matches * 2 * sizeof(Rec)
                                  : (size_t)nlit * 2 * sizeof(Rec)));
      lit_cap = (uint32_t)std::min<uint64_t>(ctx->recs.cap / sizeof(Rec), max_matches);
      retry = true;
    }
    if (ncand > cand_cap) {
      WS_HIP(ctx->cands.ensure((size_t)ncand * 5 / 4 * sizeof(Rec)));
      cand_cap = (uint32_t)(ctx->cands.cap / sizeof(Rec));
      retry = true;
    }
    if (!retry) break;
    if (attempt >= 4) return fail(WS_ERR_GPU_DEVICE, "scan buffers did not converge");
  }
  if (variant == WS_VARIANT_HASH)
    ctx->last_hits = (uint32_t)std::min<uint64_t>((uint64_t)ctx->last_hits + ctx->hcount[8], 0xFFFFFFFFull);
  uint32_t nlit = ctx->hcount[0];
  const uint32_t ncand = ctx->hcount[1];
  res->sbits = bits_for(plan.max_seg_len);
  res->pbits = bits_for(h.npat ? h.npat - 1 : 0);
  const uint32_t ibits = plan.batch ? bits_for(plan.items.size() - 1) : 0;
  if (res->sbits + res->pbits + ibits > 64)
    return fail(WS_ERR_INPUT_TOO_LARGE, "batch key does not fit 64 bits (items x length x patterns)");

  // Literal reductions for PatternSet::scan semantics.
  if (mode == WS_MODE_STANDARD && nlit > 0 && h.strategy != STRAT_NONE) {
    if (plan.batch)
      return fail(WS_ERR_INVALID_ARGUMENT, "standard mode is not available for batch scans yet");
    const size_t need = (size_t)nlit;
    WS_HIP(ctx->k0.ensure(need * 8));
    WS_HIP(ctx->k1.ensure(need * 8));
    WS_HIP(ctx->v0.ensure(need * 4));
    WS_HIP(ctx->v1.ensure(need * 4));
    WS_HIP(ctx->scratch.ensure(need * sizeof(Rec)));
    uint32_t maxlit = 0;
    for (auto& l : h.lits) maxlit = std::max<uint32_t>(maxlit, (uint32_t)l.size());
    ws_status rs = reduce_literals(ctx, h.strategy, nlit, res->sbits, res->pbits, maxlit);
    if (rs != WS_OK) return rs;
    nlit = ctx->hcount[0];
  }

  // Regex candidates: anchored walks (from prefilter hits, and from every
  // position for regexes without a selective prefix), then per-(item, regex)
  // chains.
  if (ctx->profiling) WS_HIP(hipEventRecord(ctx->ev[3], st));
  ctx->last.candidates = ncand;
  const uint32_t nscan = (uint32_t)h.rx_scan.size();
  if (ncand > 0 || nscan > 0) {
    if (nscan && plan.batch) {
      std::vector<uint64_t> pb(plan.items.size() + 1, 0);
      for (size_t k = 0; k < plan.items.size(); ++k) pb[k + 1] = pb[k] + (plan.items[k].y - plan.items[k].x) + 1;
      WS_HIP(ctx->posb.ensure(pb.size() * 8));
      WS_HIP(h2d_staged(ctx, ctx->posb.p, pb.data(), pb.size() * 8));
    }
    WalkParams w = walk_params(p, plan, ctx->items.as<uint2>(), ctx->counters.as<uint32_t>(), ctx->posb.as<uint64_t>());
    w.cands = ctx->cands.as<Rec>();
    w.ncand = ctx->counters.as<uint32_t>() + 1;
    w.cand_cap = cand_cap;
    // Raw matches: at most one per candidate, and per scanned position.
    uint64_t raw_need = (uint64_t)ncand + std::min<uint64_t>(w.npos_total * w.nscan, 1u << 20);
    for (int attempt = 0;; ++attempt) {
      if (raw_need > 0xFFFFFFF0ull) return fail(WS_ERR_INPUT_TOO_LARGE, "too many raw regex matches");
      WS_HIP(ctx->raw.ensure(std::max<uint64_t>(raw_need, 1) * sizeof(Rec)));
      w.raw = ctx->raw.as<Rec>();
      w.raw_cap = (uint32_t)raw_need;
      WS_HIP(launch_fill_u32(ctx->counters.as<uint32_t>() + 2, 1, 0, st));
      if (ncand) WS_HIP(launch_walk(ncand, st, w));
      if (nscan) WS_HIP(launch_rx_scan(st, w));
      ws_status s = read_counters(ctx);
      if (s != WS_OK) return s;
      if (ctx->hcount[2] <= raw_need) break;
      if (attempt >= 2) return fail(WS_ERR_GPU_DEVICE, "raw regex buffer did not converge");
      raw_need = ctx->hcount[2];
    }
    ws_status s = WS_OK;
    if (ctx->hcount[3] & kFlagTruncatedWalk)
      return fail(WS_ERR_INPUT_TOO_LARGE,
                  "a regex match may extend past the chunk overlap; raise the overlap to the "
                  "longest possible match");
    const uint32_t nraw = ctx->hcount[2];
    if (nraw > 0) {
      const uint32_t rbits = bits_for(h.rx.size() - 1);
      const uint32_t cbits = res->sbits + rbits + ibits;
      if (cbits > 64) return fail(WS_ERR_INPUT_TOO_LARGE, "chain key does not fit 64 bits");
      WS_HIP(ctx->k0.ensure((size_t)nraw * 8));
      WS_HIP(ctx->k1.ensure((size_t)nraw * 8));
      WS_HIP(ctx->v0.ensure((size_t)nraw * 4));
      WS_HIP(ctx->v1.ensure((size_t)nraw * 4));
      WS_HIP(launch_chain_keys(ctx->raw.as<Rec>(), nraw, nullptr, res->sbits, rbits, ctx->k0.as<uint64_t>(),
                               ctx->v0.as<uint32_t>(), st));
      ws_status s2 = sort_into(ctx, nraw, cbits, ctx->k0.as<uint64_t>(), ctx->k1.as<uint64_t>(),
                               ctx->v0.as<uint32_t>(), ctx->v1.as<uint32_t>());
      if (s2 != WS_OK) return s2;
      const uint32_t nrx = (uint32_t)h.rx.size();
      uint32_t* cin = plan.use_carried ? ctx->carried.as<uint32_t>() + (size_t)ctx->carry_phase * nrx : nullptr;
      uint32_t* cout = plan.use_carried ? ctx->carried.as<uint32_t>() + (size_t)(ctx->carry_phase ^ 1u) * nrx : nullptr;
      // The chain appends kept matches after the literal matches.
      if ((size_t)nlit + nraw > ctx->recs.cap / sizeof(Rec)) {
        // Grow while keeping the literal matches.
        DevBuf bigger;
        bigger.epoch = ctx->recs.epoch;
        WS_HIP(bigger.ensure(((size_t)nlit + nraw) * sizeof(Rec)));
        WS_HIP(hipMemcpyAsync(bigger.p, ctx->recs.p, (size_t)nlit * sizeof(Rec),
                              hipMemcpyDeviceToDevice, st));
        WS_HIP(hipStreamSynchronize(st));
        ctx->recs.release();
        ctx->recs = bigger;
      }
      WS_HIP(launch_fill_u32(ctx->counters.as<uint32_t>(), 1, nlit, st));
      WS_HIP(launch_chain_walk(ctx->k1.as<uint64_t>(), ctx->v1.as<uint32_t>(), ctx->raw.as<Rec>(),
                               nraw, nullptr, res->sbits, p->rx_maxlen, 0, cin, cout, nrx,
                               plan.seg_off, p->rx_pid, 0u, ctx->recs.as<Rec>(),
                               (uint32_t)(ctx->recs.cap / sizeof(Rec)), ctx->counters.as<uint32_t>(), st));
      if (plan.use_carried) ctx->carry_phase ^= 1u;
      s = read_counters(ctx);
      if (s != WS_OK) return s;
      nlit = ctx->hcount[0];
    }
  }
  if (ctx->profiling) WS_HIP(hipEventRecord(ctx->ev[4], st));
  const uint32_t total = nlit;
  ctx->last.raw_matches = total;
  if (total > max_matches) {
    res->count = total;
    return fail(WS_ERR_MATCH_BUFFER_OVERFLOW, "too many matches (" + std::to_string(total) +
                                                  " exceeds " + std::to_string(max_matches) + ")");
  }
  // Order by (item, start, pattern_id).
  if (total > 0) {
    WS_HIP(ctx->k0.ensure((size_t)total * 8));
    WS_HIP(ctx->k1.ensure((size_t)total * 8));
    WS_HIP(ctx->v0.ensure((size_t)total * 4));
    WS_HIP(ctx->v1.ensure((size_t)total * 4));
    res->count = total;
    ws_status s = WS_OK;
    if (plan.batch) {
      // Items end to end in id order: dense (position, pid) keys, so the
      // bucket sort applies (k_pack_tagged_dense maps positions back). An
      // item of length L takes L + 1 positions: a match may start at its end
      // (an empty match), and an empty item's matches keep their own place.
      std::vector<uint64_t> bases(plan.items.size());
      uint64_t acc = 0;
      for (size_t i = 0; i < plan.items.size(); ++i) {
        bases[i] = acc;
        acc += (uint64_t)(plan.items[i].y - plan.items[i].x) + 1;
      }
      const uint32_t gbits = bits_for(acc ? acc - 1 : 0);
      if (gbits + res->pbits > 63) return fail(WS_ERR_INPUT_TOO_LARGE, "batch key does not fit 63 bits");
      WS_HIP(ctx->bases.ensure(bases.size() * 8));
      WS_HIP(h2d_staged(ctx, ctx->bases.p, bases.data(), bases.size() * 8));
      WS_HIP(launch_order_keys_dense(ctx->recs.as<Rec>(), total, res->pbits, ctx->bases.as<uint64_t>(),
                                     ctx->k0.as<uint64_t>(), ctx->v0.as<uint32_t>(), st));
      res->dense_batch = true;
      s = order_pairs(ctx, total, gbits + res->pbits, true, res);
    } else {
      WS_HIP(launch_order_keys(ctx->recs.as<Rec>(), total, nullptr, res->sbits, res->pbits,
                               ctx->k0.as<uint64_t>(), ctx->v0.as<uint32_t>(), st));
      s = order_pairs(ctx, total, res->sbits + res->pbits + ibits, true, res);
    }
    if (s != WS_OK) return s;
  }
  res->count = total;
  return WS_OK;
}

// Reads the last run's event intervals into ctx->last; with `accumulate`
// they are added (a stream scan sums its chunks).
void finish_timings(ws_ctx* ctx, bool accumulate = false) {
  if (!ctx->profiling) return;
  (void)hipEventRecord(ctx->ev[5], ctx->stream);
  (void)hipEventSynchronize(ctx->ev[5]);
  float a = 0, b = 0, c2 = 0, d = 0, v = 0;
  (void)hipEventElapsedTime(&a, ctx->ev[1], ctx->ev[6]);
  (void)hipEventElapsedTime(&v, ctx->ev[6], ctx->ev[2]);
  if (accumulate) {
    ctx->last.prefilter_ms += a;
    ctx->last.verify_ms += v;
    ctx->last.regex_ms += b;
    ctx->last.order_ms += c2;
    ctx->last.total_ms += d;
    ctx->last.hits += ctx->last_hits;
    return;
  }
  ctx->last.verify_ms = v;
  ctx->last.hits = ctx->last_hits;
  (void)hipEventElapsedTime(&b, ctx->ev[3], ctx->ev[4]);
  (void)hipEventElapsedTime(&c2, ctx->ev[4], ctx->ev[5]);
  (void)hipEventElapsedTime(&d, ctx->ev[0], ctx->ev[5]);
  ctx->last.prefilter_ms = a;
  ctx->last.regex_ms = b;
  ctx->last.order_ms = c2;
  ctx->last.total_ms = d;
}

bool is_device_ptr(const void* p) {
  hipPointerAttribute_t attr{};
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice;
}

// Lays out one contiguous input (host or device) as a single-segment plan.
// with_prev: data[-1] exists and is the byte before the segment (a stream
// chunk after the first); the regex walks read it for their start context.
ws_status plan_single(ws_ctx* ctx, const uint8_t* data, uint64_t len, uint32_t data_mem, Plan* plan,
                      bool with_prev = false) {
  if (len > kMaxScanBytes)
    return fail(WS_ERR_INPUT_TOO_LARGE, "scan input is too large (" + std::to_string(len) + " bytes; at most " +
                                            std::to_string(kMaxScanBytes) + ")");
  const uint8_t* dptr = data;
  if (data_mem == WS_MEM_HOST) {
    // Staged from offset 16, the byte before it (with_prev) at 15.
    const size_t off = with_prev ? 16 : 0;
    WS_HIP(ctx->input.ensure(off + ((len + 15) & ~15ull) + 16));
    if (with_prev)
      WS_HIP(hipMemcpyAsync(ctx->input.as<uint8_t>() + 15, data - 1, len + 1, hipMemcpyHostToDevice, ctx->stream));
    else if (len)
      WS_HIP(hipMemcpyAsync(ctx->input.p, data, len, hipMemcpyHostToDevice, ctx->stream));
    dptr = ctx->input.as<uint8_t>() + off;
  }
  plan->ctx_before = with_prev;
  const uintptr_t a = reinterpret_cast<uintptr_t>(dptr);
  plan->base = reinterpret_cast<const uint8_t*>(a & ~(uintptr_t)15);
  const uint32_t shift = (uint32_t)(a & 15);
  plan->items = {make_uint2(shift, (uint32_t)(shift + len))};
  plan->load_end = (shift + len + 15) & ~15ull;
  plan->max_seg_len = len;
  plan->batch = false;
  return WS_OK;
}

// ---------------------------------------------------------------------------
// Device-sized runs (overlapping mode). Every pass after the filter reads its
// element count from device memory, so a run -- or every chunk of a stream --
// is enqueued without a host synchronisation. Buffers are sized to
// capacities up front; k_run_close notes any overflow (the scan is then
// redone on the checked path, whose buffers grow on demand) and appends the
// run's match count to the output offset.
// ---------------------------------------------------------------------------
// Grows the pinned staging area (reallocation makes captured graphs stale).

struct DeviceRunState {
  void* out = nullptr;       // device ws_match array
  uint64_t out_cap = 0;
  unsigned long long* total = nullptr;
  uint32_t* sticky = nullptr;
};

uint32_t bucket_bits(uint32_t keybits, uint32_t cap) {
  return std::min<uint32_t>(keybits, std::min(22u, std::max(10u, bits_for(cap) + 1)));
}

// Everything a device-sized run needs at launch time, decided (and every
// buffer sized) before any launch, so the launches can be captured.
struct RunSetup {
  uint32_t variant = 0, lane_bytes = 0, nunits = 0, grid = 0, nvb = 0;
  uint32_t lit_cap = 0, cand_cap = 0, raw_cap = 0, rec_cap = 0, max_matches = 0;
  uint64_t lslice_total = 0, cslice_total = 0;
  uint32_t sbits = 0, pbits = 0, rbits = 0, obits = 0, cbits = 0, ob = 0, cb = 0;
  uint32_t gbits = 0;  // a batch: bits of a dense position (chain and ordering keys)
  uint32_t lb = 0;       // literal-only ordering: 2^lb fixed-capacity buckets (0: not used)
  bool small_order = false;  // literal-only ordering by the one-block kernel instead
  // Two-stream runs: the slice buffers' parity and the filter's counter
  // block (null: the run counters).
  uint32_t parity = 0;
  uint32_t* fcnt = nullptr;
  size_t items_off = 0;  // this run's index among the call's runs (its segment, its stamps)
  unsigned long long* stamps = nullptr;  // the call's kernel stamps (profiling), two per run
  uint32_t* publish = nullptr;           // the call's last run: pinned host copy of the state block
  uint32_t publish_words = 0;
};

// Fixed capacities of an allocation-free call (instead of capacities that
// follow the input size).
struct RunLimits {
  uint32_t lit_cap, cand_cap;
  uint64_t slice_records;  // records of each output-slice buffer (slices + overflow area)
};

ws_status prepare_run(ws_ctx* ctx, const ws_patterns* p, const Plan& plan, const ws_tuning* tuning,
                      RunSetup* rs, const RunLimits* lim = nullptr) {
  const PatternsHost& h = p->h;
  if (plan.batch && !plan.dev_items)
    return fail(WS_ERR_INVALID_ARGUMENT, "device-sized runs take single segments or batch plans");
  rs->variant = pick_variant(p, tuning);
  if (!(supported_variants(p) & (1u << rs->variant)))
    return fail(WS_ERR_INVALID_ARGUMENT, "kernel variant not available for this pattern set");
  rs->lane_bytes = choose_lane_bytes(plan, tuning);
  const uint32_t unit = rs->variant == WS_VARIANT_HASH ? kTPB * rs->lane_bytes : ac_unit_bytes();
  if (plan.batch && rs->variant != WS_VARIANT_HASH && unit != kTPB * 64u)
    return fail(WS_ERR_GPU_DEVICE, "factor-DFA work units differ from the batch plan's");
  const uint32_t ui = unit_index(rs->variant, rs->lane_bytes);
  std::vector<uint2> none;
  rs->nunits = plan.batch ? plan.dev_nunits[ui] : plan_units(plan, unit, &none);
  if (plan.batch && !plan.dev_units[ui] && rs->nunits)
    return fail(WS_ERR_INVALID_ARGUMENT, "batch plan has no work units for this lane width");
  const uint64_t span = plan.batch ? plan.total_bytes : plan.items[0].y - plan.items[0].x;
  rs->max_matches = (uint32_t)std::min<uint64_t>(ctx->cfg.max_matches, 0xFFFFFFF0ull);
  rs->lit_cap = (uint32_t)std::min<uint64_t>(rs->max_matches, std::max<uint64_t>(1u << 16, span / 16));
  rs->cand_cap = (uint32_t)std::min<uint64_t>(0x0FFFFFFFull, std::max<uint64_t>(1u << 16, span / 64));
  if (lim) {
    rs->lit_cap = std::min(rs->max_matches, lim->lit_cap);
    rs->cand_cap = lim->cand_cap;
  }
  const uint32_t nrx = (uint32_t)h.rx.size();
  // Raw regex matches: one per candidate, plus (regexes without a selective
  // prefix) one per scanned position, up to 4 M more; a denser run overflows
  // to the checked pipeline, which sizes them exactly.
  const uint64_t scan_raw = h.rx_scan.empty() ? 0 : scan_positions(plan) * h.rx_scan.size();
  rs->raw_cap = nrx ? (uint32_t)std::min<uint64_t>(0x0FFFFFFFull, (uint64_t)rs->cand_cap +
                                                                       (lim ? 0 : std::min<uint64_t>(scan_raw, 1u << 22)))
                    : 0;
  rs->rec_cap = rs->lit_cap + rs->raw_cap;
  WS_HIP(ctx->recs.ensure((size_t)rs->rec_cap * sizeof(Rec)));
  WS_HIP(ctx->cands.ensure((size_t)rs->cand_cap * sizeof(Rec)));
  if (nrx) WS_HIP(ctx->raw.ensure((size_t)rs->raw_cap * sizeof(Rec)));
  WS_HIP(ctx->k0.ensure((size_t)rs->rec_cap * 8));
  WS_HIP(ctx->k1.ensure((size_t)rs->rec_cap * 8));
  WS_HIP(ctx->v0.ensure((size_t)rs->rec_cap * 4));
  WS_HIP(ctx->v1.ensure((size_t)rs->rec_cap * 4));
  const uint32_t per_cu = tuning && tuning->blocks_per_cu ? tuning->blocks_per_cu : 4;
  rs->grid = std::max(1u, std::min(rs->nunits, ctx->cus * per_cu));
  // Filter blocks = output slices (hash variant); the slices keep whatever
  // capacity an earlier (checked) scan grew them to.
  rs->nvb = std::max(1u, std::min((rs->nunits + 7) / 8, ctx->cus * (per_cu > 4 ? 2u : 1u)));
  if (rs->variant == WS_VARIANT_HASH) {
    const uint64_t srec = lim ? lim->slice_records : slice_records(rs->nvb);
    WS_HIP(ctx->lslices.ensure(srec * sizeof(Rec)));
    WS_HIP(ctx->cslices.ensure(srec * sizeof(Rec)));
    WS_HIP(ctx->scounts.ensure(((size_t)rs->nvb * 2 + 4) * 4 * 2));
    rs->lslice_total = ctx->lslices.cap / sizeof(Rec);
    rs->cslice_total = ctx->cslices.cap / sizeof(Rec);
  }
  rs->sbits = bits_for(plan.max_seg_len);
  rs->pbits = bits_for(h.npat ? h.npat - 1 : 0);
  rs->rbits = nrx ? bits_for(nrx - 1) : 0;
  // Ordering keys: start (one segment) or dense position (a batch), then the
  // pattern id; chain keys: item, regex, start.
  // A batch's keys are dense positions: ordering (position, pattern id),
  // chains (regex, position), each (item, regex) group contiguous.
  rs->gbits = plan.batch ? bits_for(plan.total_pos ? plan.total_pos - 1 : 0) : 0;
  rs->obits = (plan.batch ? rs->gbits : rs->sbits) + rs->pbits;
  rs->cbits = (plan.batch ? rs->gbits : rs->sbits) + rs->rbits;
  if (rs->obits > 63 || rs->cbits > 63) return fail(WS_ERR_INPUT_TOO_LARGE, "match key does not fit 63 bits");
  // Bucket counts follow the capacities up to 2^14 buckets (4 pairs per
  // bucket at 2^16 records), or 2^16 after a dense call (16 pairs per
  // bucket at a million); denser runs fall back through the overflow flag.
  const uint32_t bcap = ctx->sort_wide ? (1u << 14) : (1u << 12);
  rs->ob = bucket_bits(rs->obits, std::min(rs->rec_cap, bcap));
  rs->cb = nrx ? bucket_bits(rs->cbits, std::min(rs->raw_cap, bcap)) : 0;
  const uint32_t nb = 1u << std::max(rs->ob, rs->cb);
  WS_HIP(ensure_bucket_counts(ctx, nb));
  if (!nrx && rs->variant == WS_VARIANT_HASH) {
    // Literal-only sets: about 128 matches per bucket at the capacity, so
    // a bucket holds twice its share before the checked path takes over.
    rs->lb = std::min(rs->obits, std::min(bits_for(kLitBuckets - 1),
                                          std::max(6u, bits_for(rs->lit_cap) > 7 ? bits_for(rs->lit_cap) - 7 : 6u)));
    const uint32_t lnb = 1u << rs->lb;
    WS_HIP(ctx->lkeys.ensure((size_t)lnb * kLitBucketCap * 8));
    WS_HIP(ensure_zeroed_counts(ctx, ctx->lcnt, kLitBuckets));
    WS_HIP(ctx->lofs.ensure((size_t)kLitBuckets * 4));
    rs->small_order = ctx->lit_small && !plan.batch && rs->nvb <= kSmallOrderSlices;
  }
  if (nrx && rs->variant == WS_VARIANT_HASH) {
    // Regex sets on the hash variant: final ordering pairs (k2, v2) and their
    // counts (bcnt) are made while the chain sort uses k0..v1 and bcnt2.
    WS_HIP(ensure_bucket_counts(ctx, ctx->bcnt2, 1u << rs->cb));
    WS_HIP(ctx->k2.ensure((size_t)rs->rec_cap * 8));
    WS_HIP(ctx->v2.ensure((size_t)rs->rec_cap * 4));
  }
  WS_HIP(ctx->bcur.ensure((size_t)std::max(nb, kScanOneBlock) * 4));
  size_t need = 0, need2 = 0;
  WS_HIP(bucket_sort_pairs(nullptr, &need, nullptr, nullptr, nullptr, nullptr, rs->rec_cap, nullptr, 0,
                           1u << rs->ob, nullptr, nullptr, nullptr, ctx->stream));
  if (nrx)
    WS_HIP(bucket_sort_pairs(nullptr, &need2, nullptr, nullptr, nullptr, nullptr, rs->raw_cap, nullptr, 0,
                             1u << rs->cb, nullptr, nullptr, nullptr, ctx->stream));
  WS_HIP(ctx->tmp.ensure(std::max(need, need2)));
  if (rs->variant == WS_VARIANT_HASH) {
    // Dynamic LDS attribute of the filter kernel: set outside any capture.
    WS_HIP(prepare_hash(h.K, h.ci, rs->lane_bytes, (uint32_t)h.bloom.size(), !h.rx.empty()));
  } else {
    WS_HIP(prepare_ac(rs->variant == WS_VARIANT_DFA_LDS, h.ac_states, h.ac_classes));
  }
  return WS_OK;
}

// Launches one prepared run: stream operations only (capturable).
// part: kRunAll, or (two-stream runs, hash variant) kRunFilter (the filter
// alone, on stream2) and kRunRest (everything after it, on the stream).
enum RunPart { kRunAll = 0, kRunFilter = 1, kRunRest = 2 };

ws_status launch_run(ws_ctx* ctx, const ws_patterns* p, const Plan& plan, const RunSetup& rs, uint32_t start_off,
                     const DeviceRunState& ds, hipEvent_t* ev, RunPart part = kRunAll) {
  const PatternsHost& h = p->h;
  const hipStream_t st = ctx->stream;
  const uint32_t nrx = (uint32_t)h.rx.size();
  uint32_t* cnt = ctx->counters.as<uint32_t>();
  // This run's segment (all runs' segments were copied in one go), or a
  // batch plan's segments.
  const uint2* items = plan.dev_items ? plan.dev_items : ctx->items.as<uint2>() + rs.items_off;
  // Filter and verification.
  if (ev) WS_HIP(hipEventRecord(ev[0], st));
  // cnt[0..10] are zero here: cleared before the first run, and by
  // k_run_close after each one.
  ScanCommon c = make_common(ctx, p, plan, rs.nunits, rs.lit_cap, rs.cand_cap);
  c.items = items;
  if (plan.batch) c.units = plan.dev_units[unit_index(rs.variant, rs.lane_bytes)];
  if (ev && rs.stamps) c.stamp = rs.stamps + 2 * rs.items_off;
  RunCaps caps{};
  caps.lit_cap = rs.lit_cap;
  caps.cand_cap = rs.cand_cap;
  caps.raw_cap = rs.raw_cap;
  caps.rec_cap = rs.rec_cap;
  caps.max_matches = rs.max_matches;
  caps.lit_ovf_cap = caps.cand_ovf_cap = 0xFFFFFFFFu;
  bool keyed = false;  // ordering pairs already made (and counted) by the compaction
  SliceOut cand_slices{};  // keyed regex runs: the walk reads the candidates from the filter's slices
  const bool filter = rs.nunits > 0 && !h.fac.empty();  // no factors: no filter pass
  if (part == kRunFilter && !(filter && rs.variant == WS_VARIANT_HASH)) return WS_OK;  // nothing to filter
  if (filter && rs.variant == WS_VARIANT_HASH) {
    const HashParams hp = hash_params(p, rs.lane_bytes);
    // The filter's slices and its counters (hits, slice overflows): the
    // run's parity's own with two streams.
    uint32_t* fc = rs.fcnt ? rs.fcnt : cnt;
    uint32_t* sc = (rs.parity ? ctx->scounts2 : ctx->scounts).as<uint32_t>();
    SliceOut lso = make_slices(rs.parity ? ctx->lslices2 : ctx->lslices, rs.lslice_total, rs.nvb, sc, fc + 5);
    SliceOut cso = make_slices(rs.parity ? ctx->cslices2 : ctx->cslices, rs.cslice_total, rs.nvb, sc + rs.nvb,
                               fc + 6);
    if (part != kRunRest) {
      ScanCommon cf = c;
      cf.counters = fc;
      WS_HIP(launch_hash(h.K, h.ci, rs.nvb, part == kRunFilter ? ctx->stream2 : st, cf, hp, lso, cso, nrx > 0));
      if (part == kRunFilter) return WS_OK;
    }
    if (ev) WS_HIP(hipEventRecord(ev[1], st));
    if (rs.lb) {
      // Literal-only: slices -> fixed-capacity buckets -> ranked and packed.
      if (ev) WS_HIP(hipEventRecord(ev[2], st));
      if (ev) WS_HIP(hipEventRecord(ev[3], st));
      caps.lit_ovf_cap = lso.ovf_cap;
      caps.cand_ovf_cap = cso.ovf_cap;
      if (rs.small_order) {
        WS_HIP(launch_lit_order_small(lso, rs.nvb, rs.pbits, rs.sbits, p->flen, start_off, ds.out, ds.out_cap, cnt,
                                      rs.fcnt, caps, ds.sticky, ds.total, rs.publish, rs.publish_words, st));
      } else {
        WS_HIP(launch_lit_order(lso, rs.nvb, rs.pbits, rs.obits - rs.lb, 1u << rs.lb, ctx->lkeys.as<uint64_t>(),
                                ctx->lcnt.as<uint32_t>(), ctx->lofs.as<uint32_t>(), cnt, p->flen, start_off,
                                ds.out, ds.out_cap, ds.total, st, plan.dense));
        WS_HIP(launch_run_close(cnt, rs.fcnt, caps, ds.sticky, ds.total, rs.publish, rs.publish_words, st));
      }
      if (ev) WS_HIP(hipEventRecord(ev[4], st));
      return WS_OK;
    }
    // The compaction emits the literals' ordering pairs and bucket counts
    // (into k0/v0, or k2/v2 when a chain sort will use k0..v1) and, for a
    // batch's regex sets, the candidate records. A single segment's regex
    // walk reads the candidates from their slices instead (one launch
    // fewer per run: the 64 MiB regex chunks' step 1.64 -> 1.55 ms; a
    // batch's slices are too uneven, its step 0.63 -> 0.66 ms that way),
    // and a set without literals has no pairs to emit (their counters are
    // zero).
    const bool from_slices = nrx > 0 && !plan.batch;
    if (!from_slices || !h.lits.empty())
      WS_HIP(launch_slice_compact_keys(lso, cso, rs.nvb, cnt + 0, cnt + 1, cnt + kCounterLitSnapshot, c.lit_cap,
                                       rs.sbits, rs.pbits, (nrx ? ctx->k2 : ctx->k0).as<uint64_t>(),
                                       (nrx ? ctx->v2 : ctx->v0).as<uint32_t>(), ctx->bcnt.as<uint32_t>(),
                                       rs.obits - rs.ob, nrx && !from_slices ? c.cand_out : nullptr, c.cand_cap,
                                       st, plan.dense));
    if (from_slices) cand_slices = cso;
    keyed = true;
    caps.lit_ovf_cap = lso.ovf_cap;
    caps.cand_ovf_cap = cso.ovf_cap;
  } else if (filter) {
    AcParams ap{p->ac_table32, p->ac_table16, p->ac_cmap, p->ac_out_start, p->ac_out_list,
                p->ac_len, h.ac_states, h.ac_classes, h.ac_first_out, h.ac_maxlen};
    WS_HIP(launch_ac(rs.variant == WS_VARIANT_DFA_LDS, rs.grid, st, c, ap));
    if (ev) WS_HIP(hipEventRecord(ev[1], st));
    WS_HIP(launch_copy_u32(cnt + kCounterLitSnapshot, cnt + 0, st));
  } else if (ev) {
    WS_HIP(hipEventRecord(ev[1], st));
  }
  if (ev) WS_HIP(hipEventRecord(ev[2], st));

  // Regex candidates: anchored walks, chain ordering, greedy chains appended
  // after the literal matches.
  if (nrx) {
    WalkParams w = walk_params(p, plan, items, cnt, plan.dense.bases);
    w.cands = ctx->cands.as<Rec>();
    w.ncand = cnt + 1;
    w.cand_cap = rs.cand_cap;
    w.raw = ctx->raw.as<Rec>();
    w.raw_cap = rs.raw_cap;
    // Hash variant: the walk writes the chain pairs and counts them (bcnt2),
    // the chain walk writes the kept matches' final pairs (k2/v2, bcnt).
    uint32_t* ccnt = keyed ? ctx->bcnt2.as<uint32_t>() : ctx->bcnt.as<uint32_t>();
    const uint32_t ksbits = plan.batch ? rs.gbits : rs.sbits;  // chain key position bits
    const uint32_t gshift = plan.batch ? kGroupByRecord : rs.sbits;
    if (keyed) {
      w.ckeys = ctx->k0.as<uint64_t>();
      w.cvals = ctx->v0.as<uint32_t>();
      w.ccnt = ccnt;
      w.cshift = rs.cbits - rs.cb;
      w.sbits = ksbits;
      w.rbits = rs.rbits;
      w.cbase = plan.batch ? plan.dense.bases : nullptr;
    }
    if (cand_slices.buf)
      WS_HIP(launch_walk_slices(st, w, cand_slices, rs.nvb, cnt + 1));
    else
      WS_HIP(launch_walk(0, st, w));
    if (w.nscan) WS_HIP(launch_rx_scan(st, w));
    if (!keyed)
      WS_HIP(launch_chain_keys(ctx->raw.as<Rec>(), rs.raw_cap, cnt + 2, ksbits, rs.rbits, ctx->k0.as<uint64_t>(),
                               ctx->v0.as<uint32_t>(), st, plan.batch ? plan.dense.bases : nullptr));
    size_t have = ctx->tmp.cap;
    WS_HIP(bucket_sort_pairs(ctx->tmp.p, &have, ctx->k0.as<uint64_t>(), ctx->k1.as<uint64_t>(),
                             ctx->v0.as<uint32_t>(), ctx->v1.as<uint32_t>(), rs.raw_cap, cnt + 2, rs.cbits - rs.cb,
                             1u << rs.cb, ccnt, ctx->bcur.as<uint32_t>(), cnt + 9, st, keyed));
    // Stream chunks: run k reads half k % 2 of `carried` (the state the
    // earlier chunks left) and writes the other half.
    uint32_t* cin = plan.use_carried ? ctx->carried.as<uint32_t>() + (rs.items_off & 1) * nrx : nullptr;
    uint32_t* cout = plan.use_carried ? ctx->carried.as<uint32_t>() + ((rs.items_off + 1) & 1) * nrx : nullptr;
    OrderOut oo{};
    if (keyed) {
      oo.keys = ctx->k2.as<uint64_t>();
      oo.vals = ctx->v2.as<uint32_t>();
      oo.cnt = ctx->bcnt.as<uint32_t>();
      oo.shift = rs.obits - rs.ob;
      oo.sbits = rs.sbits;
      oo.pbits = rs.pbits;
      oo.dense = plan.dense;
    }
    WS_HIP(launch_chain_walk(ctx->k1.as<uint64_t>(), ctx->v1.as<uint32_t>(), ctx->raw.as<Rec>(), rs.raw_cap,
                             cnt + 2, gshift, p->rx_maxlen, 0, cin, cout, nrx, plan.seg_off, p->rx_pid, 0u,
                             ctx->recs.as<Rec>(), rs.rec_cap, cnt + 0, st, keyed ? &oo : nullptr));
  }
  if (ev) WS_HIP(hipEventRecord(ev[3], st));

  // Order by (start, pattern id), pack at the running output offset.
  if (keyed) {
    // Pairs in k0/v0 (literal-only sets) or k2/v2 (regex sets; k0/v0 are free
    // again after the chain walk) are scattered into the other pair.
    uint64_t* fk = (nrx ? ctx->k2 : ctx->k0).as<uint64_t>();
    uint32_t* fv = (nrx ? ctx->v2 : ctx->v0).as<uint32_t>();
    WS_HIP(bucket_order_pack(ctx->tmp.p, ctx->tmp.cap, fk, nrx ? ctx->k0.as<uint64_t>() : ctx->k1.as<uint64_t>(),
                             fv, nrx ? ctx->v0.as<uint32_t>() : ctx->v1.as<uint32_t>(), rs.rec_cap, cnt + 0,
                             rs.obits - rs.ob,
                             1u << rs.ob, ctx->bcnt.as<uint32_t>(), ctx->bcur.as<uint32_t>(), cnt + 9, rs.pbits,
                             start_off, ds.out, ds.out_cap, ds.total, st, plan.dense));
    WS_HIP(launch_run_close(cnt, rs.fcnt, caps, ds.sticky, ds.total, rs.publish, rs.publish_words, st));
    if (ev) WS_HIP(hipEventRecord(ev[4], st));
    return WS_OK;
  }
  WS_HIP(launch_order_keys(ctx->recs.as<Rec>(), rs.rec_cap, cnt + 0, rs.sbits, rs.pbits, ctx->k0.as<uint64_t>(),
                           ctx->v0.as<uint32_t>(), st, plan.dense));
  size_t have = ctx->tmp.cap;
  WS_HIP(bucket_sort_pairs(ctx->tmp.p, &have, ctx->k0.as<uint64_t>(), ctx->k1.as<uint64_t>(),
                           ctx->v0.as<uint32_t>(), ctx->v1.as<uint32_t>(), rs.rec_cap, cnt + 0, rs.obits - rs.ob,
                           1u << rs.ob, ctx->bcnt.as<uint32_t>(), ctx->bcur.as<uint32_t>(), cnt + 9, st));
  WS_HIP(launch_pack_matches(ctx->k1.as<uint64_t>(), ctx->v1.as<uint32_t>(), rs.rec_cap, cnt + 0, rs.pbits,
                             start_off, ds.out, ds.out_cap, ds.total, st, plan.dense));
  WS_HIP(launch_run_close(cnt, rs.fcnt, caps, ds.sticky, ds.total, rs.publish, rs.publish_words, st));
  if (ev) WS_HIP(hipEventRecord(ev[4], st));
  return WS_OK;
}

enum DeviceScanResult { kDeviceScanDone, kDeviceScanRedo };

// Overlapping-mode scans take the device-sized path unless the caller asks
// for the host-checked pipeline (WS_TUNING_CHECKED).
bool device_path_ok(uint32_t mode, const ws_tuning* tuning) {
  return mode == WS_MODE_OVERLAPPING && !(tuning && (tuning->flags & WS_TUNING_CHECKED));
}

// Runs the plans (one per chunk) as device-sized runs into `out` and reports
// the total. A launch sequence seen before is captured once into a hipGraph
// and replayed from then on (the steady state of repeated scans and
// streams). *outcome = kDeviceScanRedo when a buffer overflowed; nothing was
// then written to a host `out`.
// A batch plan (one run, plans[0].batch) writes ws_tagged_match records
// (rec_bytes = sizeof(ws_tagged_match)); its segments are already on the
// device.
// Two halves: device_scan_launch enqueues the call, device_scan_complete
// waits for it and reads the outcome (ws_scan_async returns between them).
struct DevicePending {
  size_t nplans = 0;
  bool lit_only = false, small = false, items_by_arg = false, replayed = false, profiling = false;
  CallItems ci{};
  size_t state_bytes = 0;
  void* dev_out = nullptr;  // where the kernels write the records
  void* out = nullptr;      // the caller's output
  uint64_t cap = 0;
  uint32_t out_mem = 0;
  size_t rec_bytes = sizeof(ws_match);
};

ws_status device_scan_launch(ws_ctx* ctx, const ws_patterns* p, const Plan* plans, size_t nplans,
                             const uint32_t* offs, const ws_tuning* tuning, void* out, uint64_t cap,
                             uint32_t out_mem, const GraphKey* key_in, size_t rec_bytes, DevicePending* dp) {
  const hipStream_t st = ctx->stream;
  const bool batch = nplans == 1 && plans[0].batch;
  // Everything sized and staged before the first launch.
  std::vector<RunSetup> setups(nplans);
  const uint64_t max_matches = ctx->cfg.max_matches;
  DeviceRunState ds;
  if (out_mem == WS_MEM_DEVICE) {
    ds.out = out;
    ds.out_cap = cap;
  } else {
    ds.out_cap = std::min<uint64_t>(cap, max_matches);
    WS_HIP(ctx->outbuf.ensure(std::max<uint64_t>(ds.out_cap, 1) * rec_bytes));
    ds.out = ctx->outbuf.p;
  }
  // The device state block: sticky flags and the running total (64 bytes),
  // then (profiling) two kernel stamps per run; one copy reads it back.
  const size_t nstamp = ctx->profiling ? nplans : 0;
  const size_t state_bytes = 64 + nstamp * 2 * sizeof(unsigned long long);
  WS_HIP(ctx->dstate.ensure(state_bytes));
  ds.sticky = ctx->dstate.as<uint32_t>();
  ds.total = reinterpret_cast<unsigned long long*>(ds.sticky + 8);
  unsigned long long* dstamps = ctx->profiling ? reinterpret_cast<unsigned long long*>(ds.sticky + 16) : nullptr;
  // Pinned staging for the runs' items when they do not fit the init
  // kernel's arguments; the state block comes back through ctx->hstate.
  const bool items_by_arg = nplans <= kCallItemsMax;
  if (!items_by_arg) WS_HIP(ensure_stage(ctx, nplans * sizeof(uint2)));
  WS_HIP(ensure_hstate(ctx, state_bytes + 4));
  // The completion flag after the state block (set by the last kernel).
  volatile uint32_t* done_flag = ctx->hstate + state_bytes / 4;
  *done_flag = 0u;
  std::atomic_thread_fence(std::memory_order_seq_cst);
  if (!batch) WS_HIP(ctx->items.ensure(nplans * sizeof(uint2)));
  // A call whose launch sequence is already captured skips the per-run setup:
  // nothing it sizes has changed since the capture (a reallocation bumps
  // the buffer epoch, which is part of the key).
  GraphEntry* ge = nullptr;
  GraphKey key;
  // Literal-only single inputs and streams may order on the one-block path
  // (prepare_run decides per run; the flag is part of the graph key).
  const bool lit_only = !batch && p->h.rx.empty() && !p->h.fac.empty() &&
                        pick_variant(p, tuning) == WS_VARIANT_HASH;
  const bool small = lit_only && ctx->lit_small;
  // Several runs on the hash variant (stream chunks): filters on a second
  // stream, one run ahead (not with per-phase profiling events).
  const bool two = nplans >= 2 && !batch && !ctx->profiling && !p->h.fac.empty() &&
                   pick_variant(p, tuning) == WS_VARIANT_HASH;
  auto lookup = [&]() {
    if (!key_in || !ctx->graphs_on) return;
    key = *key_in;
    key.epoch = ctx->epoch.load();
    key.small = small ? 1u : 0u;
    key.wide = ctx->sort_wide ? 1u : 0u;
    ge = &ctx->graphs[key];
  };
  lookup();
  const bool captured = ge && ge->exec;
  CallItems ci{};
  ci.n = items_by_arg && !batch ? (uint32_t)nplans : 0u;
  for (size_t k = 0; k < nplans; ++k) {
    if (!captured) {
      ws_status s = prepare_run(ctx, p, plans[k], tuning, &setups[k]);
      if (s != WS_OK) return s;
    }
    setups[k].items_off = k;
    setups[k].stamps = dstamps;
    if (two) {
      setups[k].parity = (uint32_t)(k & 1);
      setups[k].fcnt = ctx->counters.as<uint32_t>() + kFilterCounterBlock * (1 + (k & 1));
    }
    if (k + 1 == nplans) {
      setups[k].publish = ctx->hstate;
      setups[k].publish_words = (uint32_t)(state_bytes / 4);
    }
    if (batch)
      continue;  // the plan's segments are on the device
    if (items_by_arg)
      ci.v[k] = plans[k].items[0];
    else
      std::memcpy(ctx->hstage + k * sizeof(uint2), plans[k].items.data(), sizeof(uint2));
  }
  if (two && !captured) {
    // Odd runs' slice buffers, as large as the even runs'.
    WS_HIP(ctx->lslices2.ensure(ctx->lslices.cap));
    WS_HIP(ctx->cslices2.ensure(ctx->cslices.cap));
    WS_HIP(ctx->scounts2.ensure(ctx->scounts.cap));
  }
  if (!captured) lookup();  // the setup may have grown buffers
  if (ctx->profiling && ctx->pev.size() < 5 * nplans) {
    const size_t had = ctx->pev.size();
    ctx->pev.resize(5 * nplans, nullptr);
    for (size_t k = had; k < ctx->pev.size(); ++k) WS_HIP(hipEventCreate(&ctx->pev[k]));
  }
  // The call's init (state reset, and the runs' segments when they fit its
  // arguments) runs eagerly ahead of the runs, captured or not: a captured
  // graph then holds only the runs, and each replay gets this call's
  // segments without a copy node.
  auto launch_init = [&]() -> ws_status {
    WS_HIP(launch_call_init(ds.sticky, 10, ctx->counters.as<uint32_t>(), kCounterWords, dstamps,
                            (uint32_t)nstamp, ci, ctx->items.as<uint2>(), st));
    return WS_OK;
  };
  auto launch_all = [&]() -> ws_status {
    if (!items_by_arg)
      WS_HIP(hipMemcpyAsync(ctx->items.p, ctx->hstage, nplans * sizeof(uint2), hipMemcpyHostToDevice, st));
    if (two) {
      // Run k's filter on stream2 once run k - 2 is done (its slices and
      // counter block are free again); run k's later passes on the stream
      // once that filter is done. Run k + 1's filter then overlaps run k's
      // regex passes and ordering. (Events: fork, filter done by parity,
      // run done by parity; a wait binds to the latest record before it.)
      const hipStream_t sb = ctx->stream2;
      WS_HIP(hipEventRecord(ctx->xev[0], st));
      WS_HIP(hipStreamWaitEvent(sb, ctx->xev[0], 0));
      for (size_t k = 0; k < nplans; ++k) {
        if (k >= 2) WS_HIP(hipStreamWaitEvent(sb, ctx->xev[3 + (k & 1)], 0));
        ws_status s = launch_run(ctx, p, plans[k], setups[k], offs[k], ds, nullptr, kRunFilter);
        if (s != WS_OK) return s;
        WS_HIP(hipEventRecord(ctx->xev[1 + (k & 1)], sb));
        WS_HIP(hipStreamWaitEvent(st, ctx->xev[1 + (k & 1)], 0));
        s = launch_run(ctx, p, plans[k], setups[k], offs[k], ds, nullptr, kRunRest);
        if (s != WS_OK) return s;
        WS_HIP(hipEventRecord(ctx->xev[3 + (k & 1)], st));
      }
      return WS_OK;
    }
    for (size_t k = 0; k < nplans; ++k) {
      ws_status s = launch_run(ctx, p, plans[k], setups[k], offs[k], ds,
                               ctx->profiling ? &ctx->pev[5 * k] : nullptr);
      if (s != WS_OK) return s;
    }
    return WS_OK;
  };
  if (ctx->profiling) WS_HIP(hipEventRecord(ctx->tev[0], st));
  {
    // The previous call on this stream left the state block reset and these
    // same segments in place: nothing for the init kernel to do.
    const bool same = ctx->init_clean && items_by_arg && ctx->clean_items.n == ci.n &&
                      std::memcmp(ctx->clean_items.v, ci.v, ci.n * sizeof(uint2)) == 0 &&
                      ctx->clean_state == ctx->dstate.p && ctx->clean_state_bytes == state_bytes &&
                      ctx->clean_items_buf == ctx->items.p;
    ctx->init_clean = false;
    if (!same) {
      ws_status s = launch_init();
      if (s != WS_OK) return s;
    }
  }
  bool replayed = false;
  if (ge && ge->exec) {
    WS_HIP(hipGraphLaunch(ge->exec, st));
    replayed = true;
  } else if (ge && ge->seen >= 1 && !ge->broken) {
    // Second sighting: capture and replay from now on.
    if (ctx->graphs.size() > 64) {
      for (auto& kv : ctx->graphs)
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
      ctx->graphs.clear();
      ge = &ctx->graphs[key];
      ge->seen = 1;
    }
    hipGraph_t graph = nullptr;
    WS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    ws_status s = launch_all();
    hipError_t e = hipStreamEndCapture(st, &graph);
    hipGraphExec_t exec = nullptr;
    if (s == WS_OK && e == hipSuccess && graph) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (s != WS_OK || e != hipSuccess || !exec) {
      (void)hipGetLastError();
      ge->broken = true;
      ws_status s2 = launch_all();
      if (s2 != WS_OK) return s2;
    } else {
      ge->exec = exec;
      WS_HIP(hipGraphLaunch(exec, st));
      replayed = true;
    }
  } else {
    if (ge) ge->seen++;
    ws_status s = launch_all();
    if (s != WS_OK) return s;
  }
  if (ctx->profiling) WS_HIP(hipEventRecord(ctx->tev[1], st));
  dp->nplans = nplans;
  dp->lit_only = lit_only;
  dp->small = small;
  dp->items_by_arg = items_by_arg;
  dp->replayed = replayed;
  dp->profiling = ctx->profiling;
  dp->ci = ci;
  dp->state_bytes = state_bytes;
  dp->dev_out = ds.out;
  dp->out = out;
  dp->cap = cap;
  dp->out_mem = out_mem;
  dp->rec_bytes = rec_bytes;
  return WS_OK;
}

// Whether the launched call has finished on the device (without waiting).
bool device_scan_ready(ws_ctx* ctx, const DevicePending& dp) {
  volatile uint32_t* done_flag = ctx->hstate + dp.state_bytes / 4;
  if (*done_flag == 1u) return true;
  const hipError_t e = hipStreamQuery(ctx->stream);
  if (e == hipErrorNotReady) (void)hipGetLastError();  // not an error of a later call
  return e == hipSuccess;
}

ws_status device_scan_complete(ws_ctx* ctx, const DevicePending& dp, uint64_t* count,
                               DeviceScanResult* outcome) {
  const hipStream_t st = ctx->stream;
  *outcome = kDeviceScanDone;
  const size_t nplans = dp.nplans, state_bytes = dp.state_bytes, rec_bytes = dp.rec_bytes;
  const bool lit_only = dp.lit_only, small = dp.small, items_by_arg = dp.items_by_arg, replayed = dp.replayed;
  const CallItems& ci = dp.ci;
  const uint64_t cap = dp.cap, max_matches = ctx->cfg.max_matches;
  const uint32_t out_mem = dp.out_mem;
  void* const out = dp.out;
  volatile uint32_t* done_flag = ctx->hstate + state_bytes / 4;
  // Completion: spin briefly on the flag the last kernel sets (a shorter
  // wake-up than a blocking wait), then block on the stream. Per-phase
  // events of an eager call need the stream finished.
  bool flagged = false;
  if (!(dp.profiling && !replayed)) {
    const auto t_spin = std::chrono::steady_clock::now();
    for (uint32_t it = 0;; ++it) {
      if (*done_flag == 1u) {
        flagged = true;
        break;
      }
      if ((it & 1023u) == 0 && std::chrono::steady_clock::now() - t_spin > std::chrono::milliseconds(5)) break;
      __builtin_ia32_pause();
    }
    std::atomic_thread_fence(std::memory_order_acquire);
  }
  if (!flagged) WS_HIP(hipStreamSynchronize(st));
  // The last run's close kernel has reset the state block (whatever the
  // outcome below), and the counters; the segments stay.
  if (items_by_arg) {
    ctx->init_clean = true;
    ctx->clean_items = ci;
    ctx->clean_state = ctx->dstate.p;
    ctx->clean_state_bytes = state_bytes;
    ctx->clean_items_buf = ctx->items.p;
  }
  const uint32_t* hs = ctx->hstate;
  if (hs[kStickyRedo]) {
    if (small) ctx->lit_small = false;
    if (hs[kStickyRedo] & 64u) ctx->sort_wide = true;  // a sort bucket overflowed
    *outcome = kDeviceScanRedo;
    return WS_OK;
  }
  uint64_t tot = 0;
  std::memcpy(&tot, hs + 8, 8);
  if (lit_only) ctx->lit_small = tot <= (uint64_t)nplans * (kSmallOrder / 4);
  ctx->sort_wide = std::max<uint64_t>(tot, hs[kStickyRaw]) > ((uint64_t)nplans << 16);
  if (dp.profiling) {
    // The prefilter from its in-kernel stamps; the other phases from the
    // events between launches, which a replayed graph does not re-record
    // (then -1: unknown).
    ws_timings t{};
    const unsigned long long* stamps = reinterpret_cast<const unsigned long long*>(hs + 16);
    for (size_t k = 0; k < nplans; ++k)
      if (stamps[2 * k + 1] > stamps[2 * k])
        t.prefilter_ms += (float)((double)(stamps[2 * k + 1] - stamps[2 * k]) / ctx->wall_khz);
    if (replayed) {
      t.verify_ms = t.regex_ms = t.order_ms = -1.0f;
    } else {
      for (size_t k = 0; k < nplans; ++k) {
        hipEvent_t* ev = &ctx->pev[5 * k];
        float b = 0, c2 = 0, d = 0;
        (void)hipEventElapsedTime(&b, ev[1], ev[2]);
        (void)hipEventElapsedTime(&c2, ev[2], ev[3]);
        (void)hipEventElapsedTime(&d, ev[3], ev[4]);
        t.verify_ms += b;
        t.regex_ms += c2;
        t.order_ms += d;
      }
    }
    float all = -1.0f;
    if (hipEventQuery(ctx->tev[1]) == hipSuccess)
      (void)hipEventElapsedTime(&all, ctx->tev[0], ctx->tev[1]);
    else
      (void)hipGetLastError();  // a not-ready query must not read as a later launch error
    t.total_ms = all;
    t.hits = hs[kStickyHits];
    t.candidates = hs[kStickyCands];
    t.raw_matches = tot;
    ctx->last = t;
  }
  *count = tot;
  if (hs[kStickyError] & 1u)
    return fail(WS_ERR_INPUT_TOO_LARGE,
                "a regex match may extend past the chunk overlap; raise the overlap to the longest possible match");
  if (hs[kStickyError] & 2u || tot > cap)
    return fail(WS_ERR_MATCH_BUFFER_OVERFLOW, "too many matches (" + std::to_string(tot) + " exceeds " +
                                                  std::to_string(std::min<uint64_t>(cap, max_matches)) + ")");
  if (out_mem == WS_MEM_HOST && tot && out) {  // (no host array: timing runs)
    WS_HIP(hipMemcpyAsync(out, dp.dev_out, tot * rec_bytes, hipMemcpyDeviceToHost, st));
    WS_HIP(hipStreamSynchronize(st));
  }
  return WS_OK;
}

ws_status device_scan(ws_ctx* ctx, const ws_patterns* p, const Plan* plans, size_t nplans, const uint32_t* offs,
                      const ws_tuning* tuning, void* out, uint64_t cap, uint32_t out_mem, uint64_t* count,
                      DeviceScanResult* outcome, const GraphKey* key_in, size_t rec_bytes = sizeof(ws_match)) {
  *outcome = kDeviceScanDone;
  DevicePending dp;
  ws_status s = device_scan_launch(ctx, p, plans, nplans, offs, tuning, out, cap, out_mem, key_in, rec_bytes, &dp);
  if (s != WS_OK) return s;
  return device_scan_complete(ctx, dp, count, outcome);
}

ws_status pack_out(ws_ctx* ctx, const Result& r, uint32_t start_off, ws_match* out, uint64_t cap,
                   uint32_t out_mem, uint64_t written) {
  uint32_t n = r.count;
  if (written >= cap) return WS_OK;
  if (written + n > cap) n = (uint32_t)(cap - written);
  if (n == 0) return WS_OK;
  Result rr = r;
  for (;;) {
    if (out_mem == WS_MEM_DEVICE) {
      WS_HIP(launch_pack_matches(ctx->k1.as<uint64_t>(), ctx->v1.as<uint32_t>(), n, nullptr, rr.pbits, start_off,
                                 out + written, n, nullptr, ctx->stream));
      WS_HIP(hipStreamSynchronize(ctx->stream));
    } else {
      WS_HIP(ctx->outbuf.ensure((size_t)n * sizeof(ws_match)));
      WS_HIP(launch_pack_matches(ctx->k1.as<uint64_t>(), ctx->v1.as<uint32_t>(), n, nullptr, rr.pbits, start_off,
                                 ctx->outbuf.p, n, nullptr, ctx->stream));
      WS_HIP(hipMemcpyAsync(out + written, ctx->outbuf.p, (size_t)n * sizeof(ws_match),
                            hipMemcpyDeviceToHost, ctx->stream));
      WS_HIP(hipStreamSynchronize(ctx->stream));
    }
    bool redone = false;
    ws_status s = confirm_order(ctx, rr, &redone);
    if (s != WS_OK) return s;
    if (!redone) break;
  }
  return WS_OK;
}

// A single scan after its plan: the device-sized pipeline, and on an
// overflow (or for the checked pipeline and standard mode) the host-checked
// one. Called with ctx->mu held.
ws_status scan_checked(ws_ctx* ctx, const ws_patterns* p, const Plan& plan, uint32_t mode, const ws_tuning* tuning,
                       ws_match* out, uint64_t cap, uint32_t out_mem, uint64_t* count) {
  Result r;
  ws_status s = run_plan(ctx, p, plan, mode, tuning, &r);
  if (s != WS_OK) {
    if (s == WS_ERR_MATCH_BUFFER_OVERFLOW) *count = r.count;
    return s;
  }
  *count = r.count;
  s = pack_out(ctx, r, 0, out, cap, out_mem, 0);
  finish_timings(ctx);
  if (s != WS_OK) return s;
  if (r.count > cap)
    return fail(WS_ERR_MATCH_BUFFER_OVERFLOW, "too many matches (" + std::to_string(r.count) +
                                                  " exceeds " + std::to_string(cap) + ")");
  return WS_OK;
}

GraphKey single_key(const ws_patterns* p, const uint8_t* data, uint64_t len, const ws_tuning* tuning,
                    ws_match* out, uint64_t cap, uint32_t out_mem, bool profiling) {
  GraphKey key;
  key.pat_serial = p->serial;
  key.data = data;
  key.len = len;
  key.cap = cap;
  key.out = out_mem == WS_MEM_DEVICE ? out : nullptr;  // host output is copied after the launches
  key.out_mem = out_mem;
  key.profiling = profiling ? 1u : 0u;
  if (tuning) {
    key.variant = tuning->variant;
    key.lane = tuning->bytes_per_lane;
    key.bpc = tuning->blocks_per_cu;
  }
  return key;
}

ws_status scan_locked(ws_ctx* ctx, const ws_patterns* p, const uint8_t* data, uint64_t len, uint32_t data_mem,
                      uint32_t mode, const ws_tuning* tuning, ws_match* out, uint64_t cap, uint32_t out_mem,
                      uint64_t* count) {
  WS_HIP(hipSetDevice(ctx->dev));
  ctx->last = ws_timings{};
  Plan plan;
  ws_status s = plan_single(ctx, data, len, data_mem, &plan);
  if (s != WS_OK) return s;
  if (device_path_ok(mode, tuning)) {
    DeviceScanResult oc;
    const GraphKey key = single_key(p, data, len, tuning, out, cap, out_mem, ctx->profiling);
    const uint32_t off0 = 0;
    s = device_scan(ctx, p, &plan, 1, &off0, tuning, out, cap, out_mem, count, &oc,
                    data_mem == WS_MEM_DEVICE ? &key : nullptr);
    if (s != WS_OK || oc == kDeviceScanDone) return s;
    ctx->last = ws_timings{};  // a buffer overflowed: redo on the checked path
  }
  return scan_checked(ctx, p, plan, mode, tuning, out, cap, out_mem, count);
}

}  // namespace

// The one ws_scan_async call of a context until ws_wait collects it: either
// launched on the device-sized pipeline (launched = true until completed), or
// already finished when ws_scan_async returned.
struct ws_async_call {
  bool launched = false;
  DevicePending dp;
  // To redo the call on the checked pipeline after an overflow.
  ws_patterns* p = nullptr;  // a reference held while launched
  Plan plan;
  uint32_t mode = 0;
  bool has_tuning = false;
  ws_tuning tuning{};
  ws_match* out = nullptr;
  uint64_t cap = 0;
  uint32_t out_mem = 0;
  // The result, once known.
  ws_status status = WS_OK;
  uint64_t count = 0;
  std::string err;
};

namespace {

// Completes a launched ws_scan_async call (ctx->mu held): every other call
// on the context does this first, so its buffers are free; the result stays
// for ws_wait.
void settle_pending(ws_ctx* ctx) {
  ws_async_call* a = ctx->pending;
  if (!a || !a->launched) return;
  a->launched = false;
  (void)hipSetDevice(ctx->dev);
  DeviceScanResult oc = kDeviceScanDone;
  uint64_t n = 0;
  ws_status s = device_scan_complete(ctx, a->dp, &n, &oc);
  if (s == WS_OK && oc == kDeviceScanRedo) {
    ctx->last = ws_timings{};
    s = scan_checked(ctx, a->p, a->plan, a->mode, a->has_tuning ? &a->tuning : nullptr, a->out, a->cap,
                     a->out_mem, &n);
  }
  a->status = s;
  a->count = n;
  if (s != WS_OK) a->err = g_err;
  ws_patterns_free(a->p);
  a->p = nullptr;
}

// Pattern sources from the C ABI, names included.
ws_status read_sources(const ws_pattern* pats, size_t count, std::vector<PatternSrc>* src) {
  if (count == 0) return fail(WS_ERR_EMPTY_PATTERN_SET, "pattern set is empty. Fix: add at least one pattern");
  if (!pats) return fail(WS_ERR_INVALID_ARGUMENT, "patterns is null");
  src->assign(count, PatternSrc{});
  for (size_t i = 0; i < count; ++i) {
    PatternSrc& d = (*src)[i];
    d.kind = pats[i].kind;
    if (pats[i].len && !pats[i].bytes) return fail(WS_ERR_INVALID_ARGUMENT, "pattern bytes are null");
    d.bytes.assign(reinterpret_cast<const char*>(pats[i].bytes), pats[i].len);
    if (pats[i].name) {
      d.name = pats[i].name;
      d.has_name = true;
    }
  }
  return WS_OK;
}

// Runs host-side compilation or index work, mapping its exceptions to statuses.
template <class F>
ws_status guarded_compile(F&& f) {
  try {
    f();
  } catch (const CompileError& e) {
    return fail((ws_status)e.status, e.msg);
  } catch (const std::bad_alloc&) {
    return fail(WS_ERR_PATTERN_SET_TOO_LARGE, "out of host memory compiling the pattern set");
  }
  return WS_OK;
}

// Uploads a host-compiled set to the context's device.
ws_status make_patterns(ws_ctx* ctx, PatternsHost&& h, ws_patterns** out) {
  ws_patterns* p = new ws_patterns();
  p->serial = g_pattern_serial.fetch_add(1);
  p->h = std::move(h);
  std::lock_guard<std::mutex> lk(ctx->mu);
  settle_pending(ctx);
  p->dev = ctx->dev;
  if (hipSetDevice(ctx->dev) != hipSuccess) {
    delete p;
    return fail(WS_ERR_GPU_DEVICE, "hipSetDevice failed");
  }
  ws_status s = upload_patterns(p);
  if (s != WS_OK) {
    p->blob.release();
    delete p;
    return s;
  }
  *out = p;
  return WS_OK;
}

// Copies a serialized index to the caller's buffer (size query when out is null).
ws_status emit_index(const std::vector<uint8_t>& b, uint8_t* out, uint64_t cap, uint64_t* len) {
  *len = b.size();
  if (!out) return WS_OK;
  if (cap < b.size())
    return fail(WS_ERR_INVALID_ARGUMENT, "index buffer holds " + std::to_string(cap) + " bytes; the index needs " +
                                             std::to_string(b.size()));
  std::memcpy(out, b.data(), b.size());
  return WS_OK;
}

uint32_t index_version(uint32_t index_flags) {
  return (index_flags & WS_INDEX_REFERENCE_V3) ? kIndexVersionRef : kIndexVersion;
}

// ---------------------------------------------------------------------------
// Allocation-free scans (ws_scan_noalloc). Every workspace of one device-sized
// run is a fixed slice of `scratch`, one array per device for the life of the
// process, shared by every context on that device: calls that overlap on one
// GPU race on it. Capacities are fixed: inputs up to kNoAllocMaxBytes, up to
// kNoAllocRecords literal matches and as many regex candidates per call.
// ---------------------------------------------------------------------------
constexpr uint64_t kNoAllocMaxBytes = 16ull << 20;
constexpr uint32_t kNoAllocRecords = 1u << 16;
constexpr uint64_t kNoAllocSliceRecords = 1u << 17;  // filter output slices (each list), overflow area included
constexpr size_t kNoAllocTmpBytes = 1u << 20;        // the bucket sorts' scan storage
constexpr uint32_t kNoAllocBuckets = 1u << 16;       // sort buckets: bucket_bits() of capacities <= 2^14 is <= 16
static_assert(kNoAllocBuckets >= kScanOneBlock, "the one-block count scan reads kScanOneBlock counts");
constexpr uint32_t kNoAllocLitBucketBits = 10;       // literal-only buckets at lit_cap = 2^16 (17 - 7 bits)

// Byte offsets of the workspaces inside `scratch` (each 256-byte aligned).
struct NoAllocLayout {
  size_t recs, cands, raw, k0, k1, k2, v0, v1, v2, lslices, cslices, scounts, counters, items, dstate,
      lkeys, counts, lcnt, bcnt2, lofs, bcur, tmp, total;
};
constexpr size_t na_align(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr NoAllocLayout no_alloc_layout() {
  NoAllocLayout L{};
  size_t o = 0;
  auto take = [&o](size_t bytes) {
    const size_t at = o;
    o = na_align(o + bytes);
    return at;
  };
  const size_t rec_cap = 2ull * kNoAllocRecords;  // literal matches + regex matches
  L.recs = take(rec_cap * sizeof(Rec));
  L.cands = take(kNoAllocRecords * sizeof(Rec));
  L.raw = take(kNoAllocRecords * sizeof(Rec));
  L.k0 = take(rec_cap * 8);
  L.k1 = take(rec_cap * 8);
  L.k2 = take(rec_cap * 8);
  L.v0 = take(rec_cap * 4);
  L.v1 = take(rec_cap * 4);
  L.v2 = take(rec_cap * 4);
  L.lslices = take(kNoAllocSliceRecords * sizeof(Rec));
  L.cslices = take(kNoAllocSliceRecords * sizeof(Rec));
  L.scounts = take(64 * 1024);
  L.counters = take(64 * 4);
  L.items = take(256);
  L.dstate = take(256);
  L.lkeys = take((size_t(1) << kNoAllocLitBucketBits) * kLitBucketCap * 8);
  // The three count arrays that must start at zero, back to back: one fill.
  L.counts = o;
  L.lcnt = take((size_t)kLitBuckets * 4);
  L.bcnt2 = take((size_t)kNoAllocBuckets * 4);
  const size_t bcnt = take((size_t)kNoAllocBuckets * 4);
  (void)bcnt;
  L.lofs = take((size_t)kLitBuckets * 4);
  L.bcur = take((size_t)kNoAllocBuckets * 4);
  L.tmp = take(kNoAllocTmpBytes);
  L.total = o;
  return L;
}
constexpr NoAllocLayout kNoAllocLayout = no_alloc_layout();
constexpr size_t kScratchBytes = 20ull << 20;
static_assert(kNoAllocLayout.total <= kScratchBytes, "allocation-free workspaces exceed the scratch array");

}  // namespace

// The allocation-free path's workspaces (device memory of every GPU the
// library runs on, held until the process ends). In host code `scratch` is
// only a stand-in for the symbol: its device address comes from
// hipGetSymbolAddress, once per device (scratch_init), and only that address
// is passed to kernels and copies.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};

// With `dev` current: looks up this device's copy of `scratch` once.
ws_status scratch_init(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return fail(WS_ERR_INVALID_ARGUMENT, "device ordinal above 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return WS_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(WS_ERR_GPU_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[dev] = reinterpret_cast<uint8_t*>(ptr);
  return WS_OK;
}

// Points the context's workspaces at slices of the device's scratch array
// for one call, and restores them afterwards (the context mutex is held).
struct ScratchViews {
  std::vector<std::pair<DevBuf*, DevBuf>> saved;
  void view(DevBuf& b, uint8_t* base, size_t off, size_t bytes) {
    saved.push_back({&b, b});
    b.p = base + off;
    b.cap = bytes;
  }
  ~ScratchViews() {
    for (auto it = saved.rbegin(); it != saved.rend(); ++it) *it->first = it->second;
  }
};

}  // namespace

// A batch laid out for the device: items in source-id order (stable; batch
// output order is (source_id, start, pattern_id, end), core/src/batch.rs:159-165),
// each item's segment relative to one 16-aligned base (host items coalesced
// into the context's input buffer, each 16-byte aligned), the ids in that
// order and the dense positions (item length + 1 each).
struct BatchLayout {
  Plan plan;
  std::vector<uint64_t> ids;
  std::vector<uint64_t> bases;  // nitems + 1
};

ws_status layout_batch(ws_ctx* ctx, const ws_item* items, size_t n_items, uint32_t data_mem, BatchLayout* L) {
  // One pass checks the id order and, for device items, finds their span;
  // `order` is built only when the ids are out of order.
  bool sorted = true;
  uintptr_t lo = UINTPTR_MAX, hi = 0;
  for (size_t k = 0; k < n_items; ++k) {
    const ws_item& it = items[k];
    if (k && items[k - 1].id > it.id) sorted = false;
    if (!it.len) continue;
    if (!it.data) return fail(WS_ERR_INVALID_ARGUMENT, "item data is null");
    lo = std::min(lo, reinterpret_cast<uintptr_t>(it.data));
    hi = std::max(hi, reinterpret_cast<uintptr_t>(it.data) + (uintptr_t)it.len);
  }
  std::vector<uint32_t> order;
  if (!sorted) {
    // (id, index) pairs sort contiguously; equal ids keep their input order.
    std::vector<std::pair<uint64_t, uint32_t>> key(n_items);
    for (size_t k = 0; k < n_items; ++k) key[k] = {items[k].id, (uint32_t)k};
    std::sort(key.begin(), key.end());
    order.resize(n_items);
    for (size_t k = 0; k < n_items; ++k) order[k] = key[k].second;
  }
  auto item_at = [&](size_t k) -> const ws_item& { return items[sorted ? k : order[k]]; };
  Plan& plan = L->plan;
  plan.batch = true;
  plan.items.resize(n_items);
  L->ids.resize(n_items);
  uint64_t max_len = 0;
  if (data_mem == WS_MEM_HOST) {
    uint64_t tot = 0;
    for (size_t k = 0; k < n_items; ++k) {
      const uint64_t len = item_at(k).len;
      plan.items[k] = make_uint2((uint32_t)tot, (uint32_t)(tot + len));
      tot += (len + 15) & ~15ull;
      if (tot > kMaxScanBytes) return fail(WS_ERR_INPUT_TOO_LARGE, "coalesced batch exceeds 4 GiB - 1 MiB");
    }
    std::vector<uint8_t> host(tot, 0);
    for (size_t k = 0; k < n_items; ++k) {
      const ws_item& it = item_at(k);
      if (it.len) std::memcpy(host.data() + plan.items[k].x, it.data, it.len);
      L->ids[k] = it.id;
      max_len = std::max<uint64_t>(max_len, it.len);
    }
    WS_HIP(ctx->input.ensure(tot + 16));
    if (tot) WS_HIP(hipMemcpyAsync(ctx->input.p, host.data(), tot, hipMemcpyHostToDevice, ctx->stream));
    WS_HIP(hipStreamSynchronize(ctx->stream));
    plan.base = ctx->input.as<uint8_t>();
    plan.load_end = tot;
  } else {
    if (lo == UINTPTR_MAX) lo = hi = 0;
    const uintptr_t base = lo & ~(uintptr_t)15;
    if (hi - base > kMaxScanBytes)
      return fail(WS_ERR_INPUT_TOO_LARGE, "device batch items span more than 4 GiB - 1 MiB");
    for (size_t k = 0; k < n_items; ++k) {
      const ws_item& it = item_at(k);
      const uint64_t o = it.len ? reinterpret_cast<uintptr_t>(it.data) - base : 0;
      plan.items[k] = make_uint2((uint32_t)o, (uint32_t)(o + it.len));
      L->ids[k] = it.id;
      max_len = std::max<uint64_t>(max_len, it.len);
    }
    plan.base = reinterpret_cast<const uint8_t*>(base);
    plan.load_end = (hi - base + 15) & ~15ull;
  }
  plan.max_seg_len = max_len;
  L->bases.resize(n_items + 1);
  uint64_t acc = 0, bytes = 0;
  for (size_t k = 0; k < n_items; ++k) {
    L->bases[k] = acc;
    const uint64_t len = plan.items[k].y - plan.items[k].x;
    acc += len + 1;
    bytes += len;
  }
  L->bases[n_items] = acc;
  plan.total_bytes = bytes;
  plan.total_pos = acc;
  return WS_OK;
}

// Copies a layout's segments, dense positions, ids and the work units of
// the lane widths in `lanes` (bit k: 16 << k positions per lane) into the
// given buffers on the context's stream (through pinned staging), and points
// the layout's plan at them.
ws_status upload_batch(ws_ctx* ctx, BatchLayout& L, DevBuf& items, DevBuf& bases, DevBuf& ids, DevBuf* const* units,
                       uint32_t lanes) {
  Plan& plan = L.plan;
  const size_t n = plan.items.size();
  WS_HIP(items.ensure(n * sizeof(uint2)));
  WS_HIP(h2d_staged(ctx, items.p, plan.items.data(), n * sizeof(uint2)));
  WS_HIP(bases.ensure((n + 1) * 8));
  WS_HIP(h2d_staged(ctx, bases.p, L.bases.data(), (n + 1) * 8));
  WS_HIP(ids.ensure(n * 8));
  WS_HIP(h2d_staged(ctx, ids.p, L.ids.data(), n * 8));
  plan.dev_items = items.as<uint2>();
  plan.dense = DenseOut{bases.as<uint64_t>(), ids.as<uint64_t>(), (uint32_t)n};
  for (uint32_t k = 0; k < 3; ++k) {
    if (!(lanes & (1u << k))) continue;
    const uint32_t unit = kTPB * (16u << k);
    std::vector<uint2> runs;
    const uint32_t nunits = plan_units(plan, unit, &runs);
    DevBuf& u = *units[k];
    WS_HIP(u.ensure((size_t)nunits * sizeof(uint4) + runs.size() * sizeof(uint2) + 16));
    uint2* druns = reinterpret_cast<uint2*>(u.as<uint4>() + nunits);
    WS_HIP(h2d_staged(ctx, druns, runs.data(), runs.size() * sizeof(uint2)));
    if (nunits)
      WS_HIP(launch_expand_units(druns, (uint32_t)runs.size(), items.as<uint2>(), unit, u.as<uint4>(), nunits,
                                 ctx->stream));
    plan.dev_units[k] = u.as<uint4>();
    plan.dev_nunits[k] = nunits;
  }
  return WS_OK;
}

// Batches take the device-sized pipeline in overlapping mode (the standard
// mode's reductions run on the checked pipeline).
bool batch_device_ok(const ws_patterns* p, uint32_t mode, const ws_tuning* tuning) {
  return device_path_ok(mode, tuning) && (supported_variants(p) & (1u << pick_variant(p, tuning)));
}

// The checked pipeline over a batch layout; the tagged matches go to `out`
// (host or device memory).
ws_status batch_checked(ws_ctx* ctx, const ws_patterns* p, const BatchLayout& L, uint32_t mode,
                        const ws_tuning* tuning, ws_tagged_match* out, uint64_t cap, uint32_t out_mem,
                        uint64_t* count) {
  Result r;
  ws_status s = run_plan(ctx, p, L.plan, mode, tuning, &r);
  if (s != WS_OK) {
    if (s == WS_ERR_MATCH_BUFFER_OVERFLOW) *count = r.count;
    return s;
  }
  *count = r.count;
  const uint32_t n = (uint32_t)std::min<uint64_t>(r.count, cap);
  if (r.bucketed) {
    // The bucket sort's overflow flag came back with this synchronisation.
    WS_HIP(hipStreamSynchronize(ctx->stream));
    bool redone = false;
    s = confirm_order(ctx, r, &redone);
    if (s != WS_OK) return s;
  }
  if (n) {
    WS_HIP(ctx->ids.ensure(L.ids.size() * 8));
    WS_HIP(h2d_staged(ctx, ctx->ids.p, L.ids.data(), L.ids.size() * 8));
    void* dst = out;
    if (out_mem != WS_MEM_DEVICE) {
      WS_HIP(ctx->outbuf.ensure((size_t)n * sizeof(ws_tagged_match)));
      dst = ctx->outbuf.p;
    }
    if (r.dense_batch)
      WS_HIP(launch_pack_tagged_dense(ctx->k1.as<uint64_t>(), ctx->v1.as<uint32_t>(), n, r.pbits,
                                      ctx->bases.as<uint64_t>(), (uint32_t)L.plan.items.size(),
                                      ctx->ids.as<uint64_t>(), dst, ctx->stream));
    else
      WS_HIP(launch_pack_tagged(ctx->k1.as<uint64_t>(), ctx->v1.as<uint32_t>(), n, r.sbits, r.pbits,
                                ctx->ids.as<uint64_t>(), dst, ctx->stream));
    if (out_mem != WS_MEM_DEVICE)
      WS_HIP(hipMemcpyAsync(out, ctx->outbuf.p, (size_t)n * sizeof(ws_tagged_match), hipMemcpyDeviceToHost,
                            ctx->stream));
    WS_HIP(hipStreamSynchronize(ctx->stream));
  }
  finish_timings(ctx);
  if (r.count > cap)
    return fail(WS_ERR_MATCH_BUFFER_OVERFLOW, "too many matches (" + std::to_string(r.count) +
                                                  " exceeds " + std::to_string(cap) + ")");
  return WS_OK;
}

// Every workspace of a context.
std::vector<DevBuf*> ctx_buffers(ws_ctx* ctx) {
  return {&ctx->recs, &ctx->cands, &ctx->raw, &ctx->k0, &ctx->v0, &ctx->k1, &ctx->v1,
          &ctx->tmp, &ctx->counters, &ctx->items, &ctx->units, &ctx->input, &ctx->outbuf,
          &ctx->scratch, &ctx->carried, &ctx->ids,
          &ctx->lslices, &ctx->cslices, &ctx->scounts, &ctx->bcnt,
          &ctx->bcur, &ctx->dstate, &ctx->k2, &ctx->v2, &ctx->bcnt2, &ctx->bases,
          &ctx->lkeys, &ctx->lcnt, &ctx->lofs, &ctx->posb, &ctx->lslices2, &ctx->cslices2,
          &ctx->scounts2};
}

extern "C" {

uint32_t ws_abi_version(void) { return WS_ABI_VERSION; }

uint32_t ws_ctx_graph_count(ws_ctx* ctx) {
  if (!ctx) return 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  uint32_t n = 0;
  for (auto& kv : ctx->graphs) n += kv.second.exec ? 1u : 0u;
  return n;
}

const char* ws_last_error(void) { return g_err.c_str(); }

ws_status ws_ctx_create(int device_ordinal, const ws_config* cfg, ws_ctx** out) {
  if (!out) return fail(WS_ERR_INVALID_ARGUMENT, "out is null");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return fail(WS_ERR_NO_GPU_ADAPTER, "no GPU adapter available");
  }
  if (device_ordinal < 0 || device_ordinal >= n)
    return fail(WS_ERR_NO_GPU_ADAPTER, "device ordinal out of range");
  ws_ctx* ctx = new ws_ctx();
  ctx->dev = device_ordinal;
  for (DevBuf* b : ctx_buffers(ctx)) b->epoch = &ctx->epoch;
  if (cfg) ctx->cfg = *cfg;
  if (!ctx->cfg.max_matches) ctx->cfg.max_matches = 1000000;
  if (!ctx->cfg.chunk_size) ctx->cfg.chunk_size = 128ull << 20;
  if (!ctx->cfg.chunk_overlap) ctx->cfg.chunk_overlap = 4096;
  auto bail = [&](const char* what, hipError_t e) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    delete ctx;
    return fail(WS_ERR_GPU_DEVICE, m);
  };
  hipError_t e = hipSetDevice(device_ordinal);
  if (e != hipSuccess) return bail("hipSetDevice", e);
  hipDeviceProp_t prop{};
  e = hipGetDeviceProperties(&prop, device_ordinal);
  if (e != hipSuccess) return bail("hipGetDeviceProperties", e);
  ctx->cus = (uint32_t)prop.multiProcessorCount;
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail("hipStreamCreate", e);
  e = hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking);
  if (e != hipSuccess) return bail("hipStreamCreate", e);
  for (auto& ev : ctx->xev) {
    e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) return bail("hipEventCreate", e);
  }
  for (auto& ev : ctx->ev) {
    e = hipEventCreate(&ev);
    if (e != hipSuccess) return bail("hipEventCreate", e);
  }
  for (auto& ev : ctx->tev) {
    e = hipEventCreate(&ev);
    if (e != hipSuccess) return bail("hipEventCreate", e);
  }
  {
    int khz = 0;
    e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device_ordinal);
    if (e != hipSuccess) return bail("hipDeviceGetAttribute(wall clock rate)", e);
    ctx->wall_khz = khz > 0 ? (uint32_t)khz : 100000u;
  }
  {
    ws_status ss = scratch_init(device_ordinal);
    if (ss != WS_OK) {
      const std::string m = g_err;
      delete ctx;
      return fail(ss, m);
    }
  }
  e = hipHostMalloc(reinterpret_cast<void**>(&ctx->hcount), 32 * sizeof(uint32_t), hipHostMallocDefault);
  if (e != hipSuccess) return bail("hipHostMalloc", e);
  e = ctx->counters.ensure(kCounterWords * sizeof(uint32_t));
  if (e != hipSuccess) return bail("hipMalloc", e);
  *out = ctx;
  return WS_OK;
}

void ws_ctx_destroy(ws_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->dev);
  if (ctx->pending) {
    settle_pending(ctx);
    delete ctx->pending;
    ctx->pending = nullptr;
  }
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
  for (DevBuf* b : ctx_buffers(ctx)) b->release();
  for (auto& ev : ctx->xev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto& ev : ctx->ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto& ev : ctx->tev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto& ev : ctx->pev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto& kv : ctx->graphs)
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
  if (ctx->hstage) (void)hipHostFree(ctx->hstage);
  if (ctx->hstate) (void)hipHostFree(ctx->hstate);
  if (ctx->hcount) (void)hipHostFree(ctx->hcount);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
  delete ctx;
}

void* ws_ctx_stream(ws_ctx* ctx) { return ctx ? reinterpret_cast<void*>(ctx->stream) : nullptr; }

uint32_t ws_ctx_cu_count(ws_ctx* ctx) { return ctx ? ctx->cus : 0; }

void ws_ctx_set_profiling(ws_ctx* ctx, int on) {
  if (ctx) ctx->profiling = on != 0;
}

void ws_ctx_set_graphs(ws_ctx* ctx, int on) {
  if (ctx) ctx->graphs_on = on != 0;
}

ws_status ws_ctx_last_timings(ws_ctx* ctx, ws_timings* out) {
  if (!ctx || !out) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  *out = ctx->last;
  return WS_OK;
}

ws_status ws_patterns_compile(ws_ctx* ctx, const ws_pattern* pats, size_t count, uint32_t flags,
                              ws_patterns** out) {
  if (!ctx || !out) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  std::vector<PatternSrc> src;
  ws_status s = read_sources(pats, count, &src);
  if (s != WS_OK) return s;
  PatternsHost h;
  s = guarded_compile([&] { h = compile_patterns(src, (flags & WS_FLAG_CASE_INSENSITIVE) != 0); });
  if (s != WS_OK) return s;
  return make_patterns(ctx, std::move(h), out);
}

void ws_patterns_retain(ws_patterns* p) {
  if (p) p->refs.fetch_add(1);
}

void ws_patterns_free(ws_patterns* p) {
  if (!p) return;
  if (p->refs.fetch_sub(1) != 1) return;
  (void)hipSetDevice(p->dev);
  (void)hipDeviceSynchronize();
  p->blob.release();
  delete p;
}

static void fill_info(const PatternsHost& h, uint32_t rx_states, uint32_t variants,
                      ws_patterns_info* out) {
  ws_patterns_info i{};
  i.pattern_count = h.npat;
  i.literal_count = (uint32_t)h.lits.size();
  i.regex_count = (uint32_t)h.rx.size();
  i.factor_count = (uint32_t)h.fac.size();
  i.min_factor_len = h.min_flen;
  i.max_factor_len = h.max_flen;
  i.min_literal_len = h.min_lit_len;
  i.max_pattern_len = h.max_match_len;
  i.dfa_states = h.ac_states;
  i.dfa_classes = h.ac_classes;
  i.regex_dfa_states = rx_states;
  i.hash_window = h.K;
  i.variants_supported = variants;
  *out = i;
}

static uint32_t rx_state_total(const PatternsHost& h) {
  uint32_t rs = 0;
  for (auto& d : h.rx) rs += d.nstates;
  return rs;
}

static uint32_t host_variants(const PatternsHost& h) {
  uint32_t v = 1u << WS_VARIANT_HASH;
  if (h.ac_ok) v |= 1u << WS_VARIANT_DFA_GLOBAL;
  if (h.ac_ok && h.ac_states <= 65535 && (size_t)h.ac_states * h.ac_classes * 2 <= kAcLdsBudget)
    v |= 1u << WS_VARIANT_DFA_LDS;
  return v;
}

ws_status ws_patterns_analyze(const ws_pattern* pats, size_t count, uint32_t flags,
                              ws_patterns_info* out) {
  if (!out) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  std::vector<PatternSrc> src;
  ws_status s = read_sources(pats, count, &src);
  if (s != WS_OK) return s;
  return guarded_compile([&] {
    PatternsHost h = compile_patterns(src, (flags & WS_FLAG_CASE_INSENSITIVE) != 0);
    fill_info(h, rx_state_total(h), host_variants(h), out);
  });
}

ws_status ws_patterns_info_get(const ws_patterns* p, ws_patterns_info* out) {
  if (!p || !out) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  fill_info(p->h, p->rx_states_total, supported_variants(p), out);
  return WS_OK;
}

ws_status ws_index_build(const ws_pattern* pats, size_t count, uint32_t flags, uint32_t index_flags,
                         uint8_t* out, uint64_t cap, uint64_t* len) {
  if (!len) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  std::vector<PatternSrc> src;
  ws_status s = read_sources(pats, count, &src);
  if (s != WS_OK) return s;
  std::vector<uint8_t> b;
  s = guarded_compile([&] {
    PatternsHost h = compile_patterns(src, (flags & WS_FLAG_CASE_INSENSITIVE) != 0);
    b = index_serialize(h, index_version(index_flags));
  });
  if (s != WS_OK) return s;
  return emit_index(b, out, cap, len);
}

ws_status ws_patterns_serialize(const ws_patterns* p, uint32_t index_flags, uint8_t* out, uint64_t cap,
                                uint64_t* len) {
  if (!p || !len) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  std::vector<uint8_t> b;
  ws_status s = guarded_compile([&] { b = index_serialize(p->h, index_version(index_flags)); });
  if (s != WS_OK) return s;
  return emit_index(b, out, cap, len);
}

ws_status ws_index_inspect(const uint8_t* data, uint64_t len, ws_index_info* out, ws_patterns_info* info) {
  if (!out || (len && !data)) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  IndexHeader hd;
  ws_status s = guarded_compile([&] {
    IndexContents ic = index_parse(data, (size_t)len);
    hd = ic.hdr;
    // The rebuild must succeed too (to_pattern_set).
    PatternsHost h = compile_patterns(ic.src, hd.ci, hd.prebuilt ? &ic.pre : nullptr);
    if (info) fill_info(h, rx_state_total(h), host_variants(h), info);
  });
  if (s != WS_OK) return s;
  ws_index_info i{};
  i.version = hd.version;
  i.case_insensitive = hd.ci ? 1 : 0;
  i.pattern_count = hd.pattern_count;
  i.literal_count = hd.literal_count;
  i.regex_count = hd.regex_count;
  i.hash_window_len = hd.hash_window_len;
  i.prebuilt = hd.prebuilt ? 1 : 0;
  *out = i;
  return WS_OK;
}

ws_status ws_index_load(ws_ctx* ctx, const uint8_t* data, uint64_t len, ws_patterns** out) {
  if (!ctx || !out || (len && !data)) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  PatternsHost h;
  ws_status s = guarded_compile([&] {
    IndexContents ic = index_parse(data, (size_t)len);
    h = compile_patterns(ic.src, ic.hdr.ci, ic.hdr.prebuilt ? &ic.pre : nullptr);
  });
  if (s != WS_OK) return s;
  return make_patterns(ctx, std::move(h), out);
}

ws_status ws_index_load_file(ws_ctx* ctx, const char* path, ws_patterns** out) {
  if (!ctx || !path || !out) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f)
    return fail(WS_ERR_PATTERN_COMPILATION_FAILED, std::string("failed to read compiled index from ") + path +
                                                        ": " + std::strerror(errno) +
                                                        ". Fix: verify the file exists and is readable");
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (size < 0) {
    std::fclose(f);
    return fail(WS_ERR_PATTERN_COMPILATION_FAILED, std::string("failed to stat compiled index ") + path);
  }
  if ((uint64_t)size > kIndexMaxFileBytes) {
    std::fclose(f);
    return fail(WS_ERR_INPUT_TOO_LARGE, "compiled index is " + std::to_string(size) + " bytes; the limit is " +
                                            std::to_string(kIndexMaxFileBytes));
  }
  std::vector<uint8_t> b((size_t)size);
  const size_t got = size ? std::fread(b.data(), 1, b.size(), f) : 0;
  std::fclose(f);
  if (got != b.size())
    return fail(WS_ERR_PATTERN_COMPILATION_FAILED, std::string("short read of compiled index ") + path);
  return ws_index_load(ctx, b.data(), b.size(), out);
}

ws_status ws_index_save_file(const uint8_t* data, uint64_t len, const char* path) {
  if (!path || (len && !data)) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  FILE* f = std::fopen(path, "wb");
  if (!f)
    return fail(WS_ERR_PATTERN_COMPILATION_FAILED, std::string("failed to write compiled index to ") + path +
                                                        ": " + std::strerror(errno) +
                                                        ". Fix: verify the parent directory exists and is writable");
  const size_t put = len ? std::fwrite(data, 1, (size_t)len, f) : 0;
  const int rc = std::fclose(f);
  if (put != len || rc != 0)
    return fail(WS_ERR_PATTERN_COMPILATION_FAILED, std::string("failed to write compiled index to ") + path);
  return WS_OK;
}

ws_status ws_patterns_name(const ws_patterns* p, uint32_t id, char* buf, uint64_t cap, uint64_t* len) {
  if (!p || !len) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  if (id >= p->h.npat) return fail(WS_ERR_INVALID_ARGUMENT, "pattern id out of range");
  if (!p->h.has_name[id]) {
    *len = UINT64_MAX;
    return WS_OK;
  }
  const std::string& n = p->h.names[id];
  *len = n.size();
  if (!buf) return WS_OK;
  if (cap < n.size()) return fail(WS_ERR_INVALID_ARGUMENT, "name buffer is too small");
  std::memcpy(buf, n.data(), n.size());
  return WS_OK;
}

ws_status ws_scan(ws_ctx* ctx, const ws_patterns* p, const uint8_t* data, uint64_t len,
                  uint32_t data_mem, uint32_t mode, const ws_tuning* tuning, ws_match* out,
                  uint64_t cap, uint32_t out_mem, uint64_t* count) {
  if (!ctx || !p || !count) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  if (p->dev != ctx->dev) return fail(WS_ERR_INVALID_ARGUMENT, "pattern set was compiled for another device");
  if (len && !data) return fail(WS_ERR_INVALID_ARGUMENT, "data is null");
  if (cap && !out) return fail(WS_ERR_INVALID_ARGUMENT, "out is null");
  if (mode > WS_MODE_STANDARD) return fail(WS_ERR_INVALID_ARGUMENT, "unknown mode");
  *count = 0;
  // An empty input matches only regexes that can match the empty string.
  if (len == 0 && p->h.rx_scan.empty()) return WS_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  settle_pending(ctx);
  return scan_locked(ctx, p, data, len, data_mem, mode, tuning, out, cap, out_mem, count);
}

ws_status ws_scan_async(ws_ctx* ctx, const ws_patterns* p, const uint8_t* data, uint64_t len,
                        uint32_t data_mem, uint32_t mode, const ws_tuning* tuning, ws_match* out,
                        uint64_t cap, uint32_t out_mem) {
  if (!ctx || !p) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  if (p->dev != ctx->dev) return fail(WS_ERR_INVALID_ARGUMENT, "pattern set was compiled for another device");
  if (len && !data) return fail(WS_ERR_INVALID_ARGUMENT, "data is null");
  if (cap && !out) return fail(WS_ERR_INVALID_ARGUMENT, "out is null");
  if (mode > WS_MODE_STANDARD) return fail(WS_ERR_INVALID_ARGUMENT, "unknown mode");
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (ctx->pending)
    return fail(WS_ERR_INVALID_ARGUMENT, "this context's last ws_scan_async call has not been collected by ws_wait");
  auto* a = new ws_async_call();
  a->mode = mode;
  a->has_tuning = tuning != nullptr;
  if (tuning) a->tuning = *tuning;
  a->out = out;
  a->cap = cap;
  a->out_mem = out_mem;
  ctx->pending = a;
  if (len == 0 && p->h.rx_scan.empty()) return WS_OK;  // no matches
  if (device_path_ok(mode, tuning)) {
    WS_HIP(hipSetDevice(ctx->dev));
    ctx->last = ws_timings{};
    ws_status s = plan_single(ctx, data, len, data_mem, &a->plan);
    if (s == WS_OK) {
      const GraphKey key = single_key(p, data, len, tuning, out, cap, out_mem, ctx->profiling);
      const uint32_t off0 = 0;
      s = device_scan_launch(ctx, p, &a->plan, 1, &off0, tuning, out, cap, out_mem,
                             data_mem == WS_MEM_DEVICE ? &key : nullptr, sizeof(ws_match), &a->dp);
    }
    if (s != WS_OK) {
      ctx->pending = nullptr;
      delete a;
      return s;  // nothing is pending
    }
    ws_patterns_retain(const_cast<ws_patterns*>(p));
    a->p = const_cast<ws_patterns*>(p);
    a->launched = true;
    return WS_OK;
  }
  // Other paths synchronise between their passes: finished here.
  a->status = scan_locked(ctx, p, data, len, data_mem, mode, tuning, out, cap, out_mem, &a->count);
  if (a->status != WS_OK) a->err = g_err;
  return WS_OK;
}

int ws_poll(ws_ctx* ctx) {
  if (!ctx) return 1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ws_async_call* a = ctx->pending;
  if (!a || !a->launched) return 1;
  (void)hipSetDevice(ctx->dev);
  return device_scan_ready(ctx, a->dp) ? 1 : 0;
}

ws_status ws_wait(ws_ctx* ctx, uint64_t* count) {
  if (!ctx || !count) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  ws_async_call* a = ctx->pending;
  if (!a) return fail(WS_ERR_INVALID_ARGUMENT, "no ws_scan_async call is pending on this context");
  settle_pending(ctx);
  ctx->pending = nullptr;
  *count = a->count;
  const ws_status s = a->status;
  if (s != WS_OK) fail(s, a->err);
  delete a;
  return s;
}

ws_status ws_scan_noalloc(ws_ctx* ctx, const ws_patterns* p, const uint8_t* data, uint64_t len,
                          const ws_tuning* tuning, ws_match* out, uint64_t cap, uint64_t* count) {
  if (!ctx || !p || !count) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  if (p->dev != ctx->dev) return fail(WS_ERR_INVALID_ARGUMENT, "pattern set was compiled for another device");
  if (len && !data) return fail(WS_ERR_INVALID_ARGUMENT, "data is null");
  if (cap && !out) return fail(WS_ERR_INVALID_ARGUMENT, "out is null");
  *count = 0;
  if (len > kNoAllocMaxBytes)
    return fail(WS_ERR_INPUT_TOO_LARGE, "allocation-free scan input of " + std::to_string(len) +
                                            " bytes exceeds its maximum of " + std::to_string(kNoAllocMaxBytes) +
                                            "; use ws_scan");
  if (tuning && (tuning->flags & WS_TUNING_CHECKED))
    return fail(WS_ERR_INVALID_ARGUMENT, "the allocation-free scan has no checked pipeline");
  if (len == 0 && p->h.rx_scan.empty()) return WS_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  settle_pending(ctx);
  WS_HIP(hipSetDevice(ctx->dev));
  uint8_t* base = g_scratch_ptr[ctx->dev];
  if (!base) return fail(WS_ERR_GPU_DEVICE, "scratch array not set up for this device");
  const hipStream_t st = ctx->stream;
  Plan plan;
  ws_status s = plan_single(ctx, data, len, WS_MEM_DEVICE, &plan);
  if (s != WS_OK) return s;
  ctx->init_clean = false;
  const NoAllocLayout& L = kNoAllocLayout;
  {
    ScratchViews sv;
    sv.view(ctx->recs, base, L.recs, L.cands - L.recs);
    sv.view(ctx->cands, base, L.cands, L.raw - L.cands);
    sv.view(ctx->raw, base, L.raw, L.k0 - L.raw);
    sv.view(ctx->k0, base, L.k0, L.k1 - L.k0);
    sv.view(ctx->k1, base, L.k1, L.k2 - L.k1);
    sv.view(ctx->k2, base, L.k2, L.v0 - L.k2);
    sv.view(ctx->v0, base, L.v0, L.v1 - L.v0);
    sv.view(ctx->v1, base, L.v1, L.v2 - L.v1);
    sv.view(ctx->v2, base, L.v2, L.lslices - L.v2);
    sv.view(ctx->lslices, base, L.lslices, L.cslices - L.lslices);
    sv.view(ctx->cslices, base, L.cslices, L.scounts - L.cslices);
    sv.view(ctx->scounts, base, L.scounts, L.counters - L.scounts);
    sv.view(ctx->counters, base, L.counters, L.items - L.counters);
    sv.view(ctx->items, base, L.items, L.dstate - L.items);
    sv.view(ctx->dstate, base, L.dstate, L.lkeys - L.dstate);
    sv.view(ctx->lkeys, base, L.lkeys, L.lcnt - L.lkeys);
    sv.view(ctx->lcnt, base, L.lcnt, L.bcnt2 - L.lcnt);
    sv.view(ctx->bcnt2, base, L.bcnt2, (size_t)kNoAllocBuckets * 4);
    sv.view(ctx->bcnt, base, L.bcnt2 + (size_t)kNoAllocBuckets * 4, (size_t)kNoAllocBuckets * 4);
    sv.view(ctx->lofs, base, L.lofs, L.bcur - L.lofs);
    sv.view(ctx->bcur, base, L.bcur, L.tmp - L.bcur);
    sv.view(ctx->tmp, base, L.tmp, L.total - L.tmp);
    struct Guard {
      Guard() { g_noalloc = true; }
      ~Guard() { g_noalloc = false; }
    } guard;
    const RunLimits lim{kNoAllocRecords, kNoAllocRecords, kNoAllocSliceRecords};
    RunSetup rs;
    s = prepare_run(ctx, p, plan, tuning, &rs, &lim);
    if (s != WS_OK) {
      if (std::string(g_err).find("out of memory") != std::string::npos)
        return fail(WS_ERR_INPUT_TOO_LARGE, "the allocation-free scan's workspaces are too small for this pattern "
                                            "set and input; use ws_scan");
      return s;
    }
    // The call's first kernels: the count arrays (earlier calls on this GPU
    // may have left them otherwise), then the state block, counters and segment.
    const size_t counts_words = (L.lofs - L.counts) / 4;
    WS_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(base + L.counts), (uint32_t)counts_words, 0, st));
    CallItems ci{};
    ci.n = 1;
    ci.v[0] = plan.items[0];
    uint32_t* sticky = ctx->dstate.as<uint32_t>();
    WS_HIP(launch_call_init(sticky, 10, ctx->counters.as<uint32_t>(), kCounterLitSnapshot + 1, nullptr, 0, ci,
                            ctx->items.as<uint2>(), st));
    DeviceRunState ds;
    ds.out = out;
    ds.out_cap = cap;
    ds.sticky = sticky;
    ds.total = reinterpret_cast<unsigned long long*>(sticky + 8);
    rs.items_off = 0;
    rs.stamps = nullptr;
    rs.publish = nullptr;
    s = launch_run(ctx, p, plan, rs, 0u, ds, nullptr);
    if (s != WS_OK) return s;
    uint32_t hs[16] = {};
    WS_HIP(hipMemcpyAsync(hs, sticky, sizeof hs, hipMemcpyDeviceToHost, st));
    WS_HIP(hipStreamSynchronize(st));
    uint64_t tot = 0;
    std::memcpy(&tot, hs + 8, 8);
    if (hs[kStickyRedo])
      return fail(WS_ERR_MATCH_BUFFER_OVERFLOW,
                  "the allocation-free scan's fixed buffers overflowed (more than " + std::to_string(kNoAllocRecords) +
                      " matches or regex candidates, or matches too clustered); use ws_scan");
    *count = tot;
    if (hs[kStickyError] & 1u)
      return fail(WS_ERR_INPUT_TOO_LARGE,
                  "a regex match may extend past the chunk overlap; raise the overlap to the longest possible match");
    if ((hs[kStickyError] & 2u) || tot > cap)
      return fail(WS_ERR_MATCH_BUFFER_OVERFLOW, "too many matches (" + std::to_string(tot) + " exceeds " +
                                                    std::to_string(std::min<uint64_t>(cap, ctx->cfg.max_matches)) +
                                                    ")");
  }
  return WS_OK;
}

ws_status ws_stream_scan(ws_ctx* ctx, const ws_patterns* p, const uint8_t* data, uint64_t len,
                         uint32_t data_mem, uint64_t chunk_size, uint64_t overlap, uint32_t mode,
                         const ws_tuning* tuning, ws_match* out, uint64_t cap, uint32_t out_mem,
                         uint64_t* count) {
  if (!ctx || !p || !count) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  if (p->dev != ctx->dev) return fail(WS_ERR_INVALID_ARGUMENT, "pattern set was compiled for another device");
  if (len && !data) return fail(WS_ERR_INVALID_ARGUMENT, "data is null");
  if (cap && !out) return fail(WS_ERR_INVALID_ARGUMENT, "out is null");
  if (chunk_size == 0) return fail(WS_ERR_INVALID_ARGUMENT, "chunk_size must be positive");
  if (mode > WS_MODE_STANDARD) return fail(WS_ERR_INVALID_ARGUMENT, "unknown mode");
  *count = 0;
  if (mode == WS_MODE_STANDARD && !p->h.lits.empty() && p->h.strategy != STRAT_NONE &&
      len > chunk_size) {
    // PatternSet::scan's literal reductions are one greedy pass over the
    // whole input's overlapping literal matches, so chunks cannot be reduced
    // one at a time. When every literal fits a chunk's window (overlap >=
    // longest literal - 1) the chunks' literal matches are the whole input's,
    // and the chunked result is the whole-input scan; one ws_scan covers the
    // input (up to 4 GiB - 1 MiB), so that is how it runs.
    uint64_t maxlit = 0;
    for (auto& l : p->h.lits) maxlit = std::max<uint64_t>(maxlit, l.size());
    if (overlap + 1 < maxlit)
      return fail(WS_ERR_INVALID_ARGUMENT,
                  "standard mode needs overlap >= the longest literal - 1 (" + std::to_string(maxlit - 1) +
                      ") to reduce literal matches across chunks");
    return ws_scan(ctx, p, data, len, data_mem, mode, tuning, out, cap, out_mem, count);
  }
  if (len == 0)  // no chunks: the empty input as one scan
    return p->h.rx_scan.empty() ? WS_OK : ws_scan(ctx, p, data, 0, data_mem, mode, tuning, out, cap, out_mem, count);
  if (len > kMaxScanBytes) return fail(WS_ERR_INPUT_TOO_LARGE, "stream input is too large");
  std::lock_guard<std::mutex> lk(ctx->mu);
  settle_pending(ctx);
  WS_HIP(hipSetDevice(ctx->dev));
  ctx->last = ws_timings{};
  const uint32_t nrx = (uint32_t)p->h.rx.size();
  // Two halves: the chain state of earlier chunks, and the one being raised.
  if (nrx) WS_HIP(ctx->carried.ensure(2 * (size_t)nrx * sizeof(uint32_t)));
  if (device_path_ok(mode, tuning)) {
    // Every chunk enqueued back to back, one synchronisation at the end.
    if (nrx) WS_HIP(launch_fill_u32(ctx->carried.as<uint32_t>(), 2 * nrx, 0, ctx->stream));
    std::vector<Plan> plans;
    std::vector<uint32_t> offs;
    for (uint64_t a = 0; a < len; a += chunk_size) {
      const uint64_t b = std::min(len, a + chunk_size);
      const uint64_t e = std::min(len, b + overlap);
      Plan plan;
      ws_status s = plan_single(ctx, data + a, e - a, data_mem, &plan, a > 0);
      if (s != WS_OK) return s;
      plan.emit_end = (uint32_t)(b - a);
      plan.window_truncated = e < len;
      plan.seg_off = (uint32_t)a;
      plan.use_carried = nrx > 0;
      plans.push_back(std::move(plan));
      offs.push_back((uint32_t)a);
      if (data_mem == WS_MEM_HOST) break;  // host chunks share one staging buffer: checked path
    }
    if (data_mem != WS_MEM_HOST) {
      DeviceScanResult oc;
      GraphKey key;
      key.pat_serial = p->serial;
      key.data = data;
      key.len = len;
      key.chunk = chunk_size;
      key.overlap = overlap;
      key.cap = cap;
      key.out = out_mem == WS_MEM_DEVICE ? out : nullptr;
      key.out_mem = out_mem;
      key.profiling = ctx->profiling ? 1u : 0u;
      if (tuning) {
        key.variant = tuning->variant;
        key.lane = tuning->bytes_per_lane;
        key.bpc = tuning->blocks_per_cu;
      }
      ws_status s = device_scan(ctx, p, plans.data(), plans.size(), offs.data(), tuning, out, cap, out_mem, count,
                                &oc, &key);
      if (s != WS_OK || oc == kDeviceScanDone) return s;
      ctx->last = ws_timings{};
    }
  }
  if (nrx) WS_HIP(launch_fill_u32(ctx->carried.as<uint32_t>(), 2 * nrx, 0, ctx->stream));
  ctx->carry_phase = 0;
  uint64_t total = 0;
  for (uint64_t a = 0; a < len; a += chunk_size) {
    const uint64_t b = std::min(len, a + chunk_size);
    const uint64_t e = std::min(len, b + overlap);
    Plan plan;
    ws_status s = plan_single(ctx, data + a, e - a, data_mem, &plan, a > 0);
    if (s != WS_OK) return s;
    plan.emit_end = (uint32_t)(b - a);
    plan.window_truncated = e < len;
    plan.seg_off = (uint32_t)a;
    plan.use_carried = nrx > 0;
    Result r;
    s = run_plan(ctx, p, plan, mode, tuning, &r);
    if (s != WS_OK) {
      if (s == WS_ERR_MATCH_BUFFER_OVERFLOW) *count = total + r.count;
      return s;
    }
    s = pack_out(ctx, r, (uint32_t)a, out, cap, out_mem, total);
    if (s != WS_OK) return s;
    finish_timings(ctx, true);
    total += r.count;
  }
  *count = total;
  if (total > cap)
    return fail(WS_ERR_MATCH_BUFFER_OVERFLOW, "too many matches (" + std::to_string(total) +
                                                  " exceeds " + std::to_string(cap) + ")");
  return WS_OK;
}

// The routing table: prefilter variant, lane width, blocks per CU, and the
// pipeline (device-sized, or host-checked: WS_TUNING_CHECKED). Measured on
// the 20 request pattern sets: the checked pipeline is about 30% faster on
// requests of a few KiB with regexes (its passes are sized from host-known
// counts; the device-sized one launches fixed grids) and twice as fast on
// 256 MiB requests whose regex candidates overflow the device-sized buffers;
// the device-sized one wins elsewhere. Two blocks per CU help some mid-size
// requests by a few percent.
static const ws_tuning kConfigTable[WS_ROUTER_CONFIGS] = {
    {WS_VARIANT_HASH, 16, 0, 0},                  {WS_VARIANT_HASH, 32, 0, 0},
    {WS_VARIANT_HASH, 64, 0, 0},                  {WS_VARIANT_HASH, 16, 0, WS_TUNING_CHECKED},
    {WS_VARIANT_HASH, 32, 0, WS_TUNING_CHECKED},  {WS_VARIANT_HASH, 64, 0, WS_TUNING_CHECKED},
    {WS_VARIANT_HASH, 16, 8, 0},                  {WS_VARIANT_HASH, 32, 8, 0},
    {WS_VARIANT_HASH, 64, 8, 0},                  {WS_VARIANT_DFA_GLOBAL, 0, 4, 0},
    {WS_VARIANT_DFA_GLOBAL, 0, 8, 0},             {WS_VARIANT_DFA_GLOBAL, 0, 8, WS_TUNING_CHECKED},
    {WS_VARIANT_DFA_LDS, 0, 2, 0},                {WS_VARIANT_DFA_LDS, 0, 4, 0},
    {WS_VARIANT_DFA_LDS, 0, 8, 0},                {WS_VARIANT_DFA_LDS, 0, 4, WS_TUNING_CHECKED},
};

ws_status ws_config_tuning(uint32_t index, ws_tuning* out) {
  if (!out || index >= WS_ROUTER_CONFIGS) return fail(WS_ERR_INVALID_ARGUMENT, "bad configuration index");
  *out = kConfigTable[index];
  return WS_OK;
}

ws_status ws_router_features(const ws_patterns_info* info, uint64_t bytes, float matches_per_mib,
                             float out[8]) {
  if (!info || !out) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  out[0] = std::log2((float)bytes + 1.f);
  out[1] = std::log2((float)std::max(1u, info->pattern_count));
  out[2] = std::log2((float)info->dfa_states + 1.f);
  out[3] = (float)info->dfa_classes;
  out[4] = info->pattern_count ? (float)info->literal_count / (float)info->pattern_count : 0.f;
  out[5] = (float)info->min_literal_len;
  out[6] = std::log2(1.f + std::max(0.f, matches_per_mib));
  out[7] = 0.f;
  return WS_OK;
}

ws_status ws_autotune(ws_ctx* ctx, const ws_patterns* p, const uint8_t* data, uint64_t len,
                      uint32_t data_mem, uint32_t mode, uint32_t repeats, uint32_t configs, float* times_ms,
                      float* mask, uint32_t* best) {
  static_assert(WS_ROUTER_CONFIGS <= 32, "configuration bit mask");
  if (!ctx || !p || !times_ms || !mask || !best || !data || len == 0)
    return fail(WS_ERR_INVALID_ARGUMENT, "bad arguments");
  if (p->dev != ctx->dev) return fail(WS_ERR_INVALID_ARGUMENT, "pattern set was compiled for another device");
  std::lock_guard<std::mutex> lk(ctx->mu);
  settle_pending(ctx);
  WS_HIP(hipSetDevice(ctx->dev));
  const uint32_t sup = supported_variants(p);
  int bi = -1;
  for (uint32_t i = 0; i < WS_ROUTER_CONFIGS; ++i) {
    times_ms[i] = 0.f;
    const bool selected = configs == 0 || ((configs >> i) & 1u);
    mask[i] = selected && ((sup >> kConfigTable[i].variant) & 1u) ? 1.f : 0.f;
    if (mask[i] == 0.f) continue;
    float tbest = 0.f;
    for (uint32_t r = 0; r < std::max(1u, repeats); ++r) {
      Plan plan;
      ws_status s = plan_single(ctx, data, len, data_mem, &plan);
      if (s != WS_OK) return s;
      WS_HIP(hipStreamSynchronize(ctx->stream));
      const auto t0 = std::chrono::steady_clock::now();
      if (device_path_ok(mode, &kConfigTable[i])) {
        // The path a scan of this request takes: device-sized, eager.
        DeviceScanResult oc;
        uint64_t n = 0;
        const uint32_t off0 = 0;
        s = device_scan(ctx, p, &plan, 1, &off0, &kConfigTable[i], nullptr, ctx->cfg.max_matches, WS_MEM_HOST, &n,
                        &oc, nullptr);
        if (s != WS_OK && s != WS_ERR_MATCH_BUFFER_OVERFLOW) return s;
        if (oc == kDeviceScanDone) {
          const float ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
          if (r == 0 || ms < tbest) tbest = ms;
          continue;
        }
      }
      Result res;
      s = run_plan(ctx, p, plan, mode, &kConfigTable[i], &res);
      if (s != WS_OK) return s;
      if (res.count) {
        WS_HIP(ctx->outbuf.ensure((size_t)res.count * sizeof(ws_match)));
        WS_HIP(launch_pack_matches(ctx->k1.as<uint64_t>(), ctx->v1.as<uint32_t>(), res.count, nullptr, res.pbits,
                                   0, ctx->outbuf.p, res.count, nullptr, ctx->stream));
      }
      WS_HIP(hipStreamSynchronize(ctx->stream));
      bool redone = false;
      s = confirm_order(ctx, res, &redone);  // an overflowing bucket costs the radix sort too
      if (s != WS_OK) return s;
      if (redone) WS_HIP(hipStreamSynchronize(ctx->stream));
      const float ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
      if (r == 0 || ms < tbest) tbest = ms;
    }
    times_ms[i] = tbest;
    if (bi < 0 || tbest < times_ms[bi]) bi = (int)i;
  }
  *best = bi < 0 ? 0u : (uint32_t)bi;
  return WS_OK;
}

ws_status ws_scan_batch(ws_ctx* ctx, const ws_patterns* p, const ws_item* items, size_t n_items,
                        uint32_t data_mem, uint32_t mode, const ws_tuning* tuning,
                        ws_tagged_match* out, uint64_t cap, uint64_t* count) {
  if (!ctx || !p || !count) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  if (p->dev != ctx->dev) return fail(WS_ERR_INVALID_ARGUMENT, "pattern set was compiled for another device");
  if (n_items && !items) return fail(WS_ERR_INVALID_ARGUMENT, "items is null");
  if (cap && !out) return fail(WS_ERR_INVALID_ARGUMENT, "out is null");
  if (mode > WS_MODE_STANDARD) return fail(WS_ERR_INVALID_ARGUMENT, "unknown mode");
  *count = 0;
  if (n_items == 0) return WS_OK;
  if (n_items > 0x7FFFFFFFull) return fail(WS_ERR_INPUT_TOO_LARGE, "too many batch items");
  std::lock_guard<std::mutex> lk(ctx->mu);
  settle_pending(ctx);
  WS_HIP(hipSetDevice(ctx->dev));
  ctx->last = ws_timings{};
  BatchLayout L;
  ws_status s = layout_batch(ctx, items, n_items, data_mem, &L);
  if (s != WS_OK) return s;
  if (batch_device_ok(p, mode, tuning)) {
    // Device-sized: segments, positions, ids and this lane width's work
    // units staged into the context's buffers, one synchronisation.
    ctx->stage_used = 0;
    ctx->init_clean = false;
    DevBuf* units[3] = {&ctx->units, &ctx->units, &ctx->units};
    s = upload_batch(ctx, L, ctx->items, ctx->bases, ctx->ids, units,
                     1u << unit_index(pick_variant(p, tuning), choose_lane_bytes(L.plan, tuning)));
    if (s != WS_OK) return s;
    DeviceScanResult oc;
    const uint32_t off0 = 0;
    s = device_scan(ctx, p, &L.plan, 1, &off0, tuning, out, cap, WS_MEM_HOST, count, &oc, nullptr,
                    sizeof(ws_tagged_match));
    if (s != WS_OK || oc == kDeviceScanDone) return s;
    ctx->last = ws_timings{};  // a buffer overflowed: redo on the checked path
  }
  return batch_checked(ctx, p, L, mode, tuning, out, cap, WS_MEM_HOST, count);
}

// ---------------------------------------------------------------------------
// Batch plans: a batch laid out once, scanned many times.
// ---------------------------------------------------------------------------
struct ws_batch_plan {
  int dev = 0;
  uint64_t serial = 0;
  BatchLayout L;
  DevBuf items, bases, ids, units[3];
};

ws_status ws_batch_plan_create(ws_ctx* ctx, const ws_item* items, size_t n_items, ws_batch_plan** out) {
  if (!ctx || !out) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (n_items && !items) return fail(WS_ERR_INVALID_ARGUMENT, "items is null");
  if (n_items == 0) return fail(WS_ERR_INVALID_ARGUMENT, "a batch plan needs at least one item");
  if (n_items > 0x7FFFFFFFull) return fail(WS_ERR_INPUT_TOO_LARGE, "too many batch items");
  std::lock_guard<std::mutex> lk(ctx->mu);
  settle_pending(ctx);
  WS_HIP(hipSetDevice(ctx->dev));
  auto* bp = new ws_batch_plan();
  bp->dev = ctx->dev;
  bp->serial = g_pattern_serial.fetch_add(1);
  ws_status s = layout_batch(ctx, items, n_items, WS_MEM_DEVICE, &bp->L);
  if (s == WS_OK) {
    ctx->stage_used = 0;
    ctx->init_clean = false;
    DevBuf* units[3] = {&bp->units[0], &bp->units[1], &bp->units[2]};
    s = upload_batch(ctx, bp->L, bp->items, bp->bases, bp->ids, units, 7u);
  }
  if (s == WS_OK) {
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) s = fail(WS_ERR_GPU_DEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  }
  if (s != WS_OK) {
    for (DevBuf* b : {&bp->items, &bp->bases, &bp->ids, &bp->units[0], &bp->units[1], &bp->units[2]}) b->release();
    delete bp;
    return s;
  }
  *out = bp;
  return WS_OK;
}

void ws_batch_plan_destroy(ws_batch_plan* bp) {
  if (!bp) return;
  (void)hipSetDevice(bp->dev);
  (void)hipDeviceSynchronize();  // scans of other contexts may still read the plan's buffers
  for (DevBuf* b : {&bp->items, &bp->bases, &bp->ids, &bp->units[0], &bp->units[1], &bp->units[2]}) b->release();
  delete bp;
}

ws_status ws_scan_batch_planned(ws_ctx* ctx, const ws_patterns* p, const ws_batch_plan* bp, uint32_t mode,
                                const ws_tuning* tuning, ws_tagged_match* out, uint64_t cap, uint32_t out_mem,
                                uint64_t* count) {
  if (!ctx || !p || !bp || !count) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  if (p->dev != ctx->dev || bp->dev != ctx->dev)
    return fail(WS_ERR_INVALID_ARGUMENT, "pattern set or batch plan belongs to another device");
  if (cap && !out) return fail(WS_ERR_INVALID_ARGUMENT, "out is null");
  if (mode > WS_MODE_STANDARD) return fail(WS_ERR_INVALID_ARGUMENT, "unknown mode");
  *count = 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  settle_pending(ctx);
  WS_HIP(hipSetDevice(ctx->dev));
  ctx->last = ws_timings{};
  if (batch_device_ok(p, mode, tuning)) {
    // A repeated call (same patterns, plan, output and tuning) replays a
    // captured graph.
    GraphKey key;
    key.pat_serial = p->serial;
    key.batch = bp->serial;
    key.cap = cap;
    key.out = out_mem == WS_MEM_DEVICE ? out : nullptr;
    key.out_mem = out_mem;
    key.profiling = ctx->profiling ? 1u : 0u;
    if (tuning) {
      key.variant = tuning->variant;
      key.lane = tuning->bytes_per_lane;
      key.bpc = tuning->blocks_per_cu;
    }
    DeviceScanResult oc;
    const uint32_t off0 = 0;
    ws_status s = device_scan(ctx, p, &bp->L.plan, 1, &off0, tuning, out, cap, out_mem, count, &oc, &key,
                              sizeof(ws_tagged_match));
    if (s != WS_OK || oc == kDeviceScanDone) return s;
    ctx->last = ws_timings{};
  }
  return batch_checked(ctx, p, bp->L, mode, tuning, out, cap, out_mem, count);
}

// ---------------------------------------------------------------------------
// Hot-swappable pattern set: the active set behind a mutex; every holder
// (the swap slot, each scan in flight) owns a reference.
// ---------------------------------------------------------------------------
struct ws_hotswap {
  std::mutex mu;
  ws_patterns* cur = nullptr;
};

ws_status ws_hotswap_create(ws_patterns* initial, ws_hotswap** out) {
  if (!initial || !out) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  ws_patterns_retain(initial);
  auto* hs = new ws_hotswap();
  hs->cur = initial;
  *out = hs;
  return WS_OK;
}

void ws_hotswap_destroy(ws_hotswap* hs) {
  if (!hs) return;
  ws_patterns_free(hs->cur);
  delete hs;
}

int ws_hotswap_swap(ws_hotswap* hs, ws_patterns* next) {
  if (!hs || !next) return 0;
  ws_patterns* old = nullptr;
  {
    std::lock_guard<std::mutex> lk(hs->mu);
    if (hs->cur == next) return 0;
    ws_patterns_retain(next);
    old = hs->cur;
    hs->cur = next;
  }
  ws_patterns_free(old);
  return 1;
}

ws_status ws_hotswap_scan(ws_ctx* ctx, ws_hotswap* hs, const uint8_t* data, uint64_t len, uint32_t data_mem,
                          uint32_t mode, const ws_tuning* tuning, ws_match* out, uint64_t cap, uint32_t out_mem,
                          uint64_t* count) {
  if (!hs) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  ws_patterns* p = nullptr;
  {
    std::lock_guard<std::mutex> lk(hs->mu);
    p = hs->cur;
    ws_patterns_retain(p);
  }
  const ws_status s = ws_scan(ctx, p, data, len, data_mem, mode, tuning, out, cap, out_mem, count);
  ws_patterns_free(p);
  return s;
}

// ---------------------------------------------------------------------------
// Incremental stream scanner. The window (carried tail + new chunk) lives in
// one of two device buffers; each feed builds the next window in the other
// buffer and scans it with ws_scan (so repeated window shapes replay a graph).
// ---------------------------------------------------------------------------
struct ws_stream_scanner {
  ws_ctx* ctx = nullptr;
  const ws_patterns* p = nullptr;
  uint32_t mode = WS_MODE_STANDARD;
  uint64_t keep = 0;         // longest literal: the carried tail's maximum length
  DevBuf win[2];
  int cur = 0;               // buffer holding the current window
  uint64_t win_len = 0;      // its length
  uint64_t overlap = 0;      // its carried tail (the last `overlap` bytes)
  uint64_t processed = 0;    // bytes fed so far
  std::vector<ws_match> tmp; // the window's matches
};

ws_status ws_stream_scanner_create(ws_ctx* ctx, const ws_patterns* p, uint32_t mode, ws_stream_scanner** out) {
  if (!ctx || !p || !out) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  if (mode > WS_MODE_STANDARD) return fail(WS_ERR_INVALID_ARGUMENT, "unknown mode");
  if (p->dev != ctx->dev) return fail(WS_ERR_INVALID_ARGUMENT, "pattern set was compiled for another device");
  auto* sc = new ws_stream_scanner();
  sc->ctx = ctx;
  ws_patterns_retain(const_cast<ws_patterns*>(p));
  sc->p = p;
  sc->mode = mode;
  for (const auto& l : p->h.lits) sc->keep = std::max<uint64_t>(sc->keep, l.size());
  *out = sc;
  return WS_OK;
}

void ws_stream_scanner_destroy(ws_stream_scanner* sc) {
  if (!sc) return;
  {
    std::lock_guard<std::mutex> lk(sc->ctx->mu);
    settle_pending(sc->ctx);
    (void)hipSetDevice(sc->ctx->dev);
    (void)hipStreamSynchronize(sc->ctx->stream);
    sc->win[0].release();
    sc->win[1].release();
  }
  ws_patterns_free(const_cast<ws_patterns*>(sc->p));
  delete sc;
}

uint64_t ws_stream_scanner_processed(const ws_stream_scanner* sc) { return sc ? sc->processed : 0; }

ws_status ws_stream_scanner_feed(ws_stream_scanner* sc, const uint8_t* chunk, uint64_t len, uint32_t data_mem,
                                 ws_match* out, uint64_t cap, uint64_t* count) {
  if (!sc || !count) return fail(WS_ERR_INVALID_ARGUMENT, "null argument");
  if (len && !chunk) return fail(WS_ERR_INVALID_ARGUMENT, "chunk is null");
  if (cap && !out) return fail(WS_ERR_INVALID_ARGUMENT, "out is null");
  *count = 0;
  ws_ctx* ctx = sc->ctx;
  const uint64_t combined = sc->overlap + len;
  if (combined > kMaxScanBytes)
    return fail(WS_ERR_INPUT_TOO_LARGE, "stream window of " + std::to_string(combined) + " bytes exceeds 4 GiB - 1 MiB");
  if (sc->processed + len > 0xFFFFFFFFull)
    return fail(WS_ERR_INPUT_TOO_LARGE, "stream offset " + std::to_string(sc->processed + len) + " exceeds 4 GiB");
  const int nxt = 1 - sc->cur;
  {
    // The next window: the current one's carried tail, then the chunk.
    std::lock_guard<std::mutex> lk(ctx->mu);
    settle_pending(ctx);
    WS_HIP(hipSetDevice(ctx->dev));
    WS_HIP(sc->win[nxt].ensure(((combined + 15) & ~15ull) + 16));
    if (sc->overlap)
      WS_HIP(hipMemcpyAsync(sc->win[nxt].p, sc->win[sc->cur].as<uint8_t>() + (sc->win_len - sc->overlap),
                            sc->overlap, hipMemcpyDeviceToDevice, ctx->stream));
    if (len)
      WS_HIP(hipMemcpyAsync(sc->win[nxt].as<uint8_t>() + sc->overlap, chunk, len,
                            data_mem == WS_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                            ctx->stream));
    WS_HIP(hipStreamSynchronize(ctx->stream));
  }
  // The window's matches, in order (start, pattern id, end).
  const uint64_t mcap = std::max<uint64_t>(1024, std::min<uint64_t>(ctx->cfg.max_matches, combined + 16));
  if (sc->tmp.size() < mcap) sc->tmp.resize(mcap);
  uint64_t n = 0;
  ws_status s = ws_scan(ctx, sc->p, sc->win[nxt].as<uint8_t>(), combined, WS_MEM_DEVICE, sc->mode, nullptr,
                        sc->tmp.data(), sc->tmp.size(), WS_MEM_HOST, &n);
  if (s == WS_ERR_MATCH_BUFFER_OVERFLOW && n <= ctx->cfg.max_matches) {
    sc->tmp.resize(n);
    s = ws_scan(ctx, sc->p, sc->win[nxt].as<uint8_t>(), combined, WS_MEM_DEVICE, sc->mode, nullptr,
                sc->tmp.data(), sc->tmp.size(), WS_MEM_HOST, &n);
  }
  if (s != WS_OK) return s;
  // New matches end past the carried tail; offsets become stream offsets.
  const uint64_t base = sc->processed - sc->overlap;
  uint64_t kept = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (sc->tmp[i].end > sc->overlap) ++kept;
  *count = kept;
  if (kept > cap)
    return fail(WS_ERR_MATCH_BUFFER_OVERFLOW, "too many matches (" + std::to_string(kept) + " exceeds " +
                                                  std::to_string(cap) + ")");
  uint64_t k = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const ws_match& m = sc->tmp[i];
    if (m.end <= sc->overlap) continue;
    out[k++] = ws_match{m.pattern_id, (uint32_t)(m.start + base), (uint32_t)(m.end + base)};
  }
  sc->cur = nxt;
  sc->win_len = combined;
  sc->overlap = std::min(sc->keep, combined);
  sc->processed += len;
  return WS_OK;
}

ws_status ws_stream_scanner_finish(ws_stream_scanner* sc, ws_match* out, uint64_t cap, uint64_t* count) {
  return ws_stream_scanner_feed(sc, nullptr, 0, WS_MEM_HOST, out, cap, count);
}

}  // extern "C"
