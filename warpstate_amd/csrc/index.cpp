// Compiled pattern index reader and writer (WPSIDX01, versions 1-4).
#include "index.hpp"

#include <algorithm>
#include <cstring>
#include <string>

#include "../../include/warpstate.h"

namespace ws {
namespace {

const uint8_t kMagic[8] = {'W', 'P', 'S', 'I', 'D', 'X', '0', '1'};  // mod.rs:17
constexpr uint32_t kNoName = 0xFFFFFFFFu;                              // builder.rs:128-140

[[noreturn]] void corrupt(const std::string& what) {
  throw CompileError{WS_ERR_PATTERN_COMPILATION_FAILED,
                     "compiled index " + what + ". Fix: rebuild the index"};
}

// Strict UTF-8 (what Rust's String::from_utf8 accepts): no overlong forms,
// no surrogates, nothing above U+10FFFF.
bool valid_utf8(const std::string& s) {
  const uint8_t* p = reinterpret_cast<const uint8_t*>(s.data());
  size_t i = 0, n = s.size();
  while (i < n) {
    const uint8_t c = p[i];
    if (c < 0x80) { ++i; continue; }
    size_t k;
    uint32_t lo = 0x80, hi = 0xBF;
    if (c >= 0xC2 && c <= 0xDF) k = 1;
    else if (c >= 0xE0 && c <= 0xEF) {
      k = 2;
      if (c == 0xE0) lo = 0xA0;
      if (c == 0xED) hi = 0x9F;
    } else if (c >= 0xF0 && c <= 0xF4) {
      k = 3;
      if (c == 0xF0) lo = 0x90;
      if (c == 0xF4) hi = 0x8F;
    } else {
      return false;
    }
    if (n - i <= k) return false;
    if (p[i + 1] < lo || p[i + 1] > hi) return false;
    for (size_t t = 2; t <= k; ++t)
      if (p[i + t] < 0x80 || p[i + t] > 0xBF) return false;
    i += k + 1;
  }
  return true;
}

// --------------------------------------------------------------------------
// Writer
// --------------------------------------------------------------------------

struct Out {
  std::vector<uint8_t> b;
  void u32(uint32_t v) {
    for (int s = 0; s < 32; s += 8) b.push_back((uint8_t)(v >> s));
  }
  static uint32_t fit(uint64_t v, const char* what) {
    if (v > 0xFFFFFFFFull)
      throw CompileError{WS_ERR_PATTERN_SET_TOO_LARGE,
                         std::string(what) +
                             " exceeds the on-disk 32-bit format. Fix: split the pattern set into smaller indexes"};
    return (uint32_t)v;
  }
  void bytes(const void* p, size_t n, const char* what) {
    u32(fit(n, what));
    raw(p, n);
  }
  void raw(const void* p, size_t n) {
    const uint8_t* q = static_cast<const uint8_t*>(p);
    b.insert(b.end(), q, q + n);
  }
  // The file is little-endian, as is every host this library runs on.
  void vec(const std::vector<uint32_t>& v, const char* what) {
    u32(fit(v.size(), what));
    raw(v.data(), v.size() * 4);
  }
};
static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "the index format is little-endian");

// literal_prefilter_hash (core/src/pattern/compiler.rs:343-350): FNV-1a over
// the first `window` bytes.
uint32_t fnv1a_window(const std::string& s, uint32_t window) {
  uint32_t h = 2166136261u;
  const size_t n = std::min<size_t>(s.size(), std::max<uint32_t>(window, 1));
  for (size_t i = 0; i < n; ++i) {
    h ^= (uint8_t)s[i];
    h *= 16777619u;
  }
  return h;
}

// LiteralPrefilterTable::build (core/src/literal_prefilter.rs:23-88): literals
// grouped by prefix length min(len, window) (at least 1); each group owns
// next_pow2(2 * size) buckets, filled bucket by bucket in literal order.
void write_prefilter_table(Out& o, const std::vector<std::string>& lits, uint32_t window) {
  std::vector<std::vector<uint32_t>> groups(window);
  std::vector<uint32_t> hash(lits.size());
  for (uint32_t i = 0; i < lits.size(); ++i) {
    hash[i] = fnv1a_window(lits[i], window);
    const uint32_t pl = (uint32_t)std::max<size_t>(1, std::min<size_t>(lits[i].size(), window));
    groups[pl - 1].push_back(i);
  }
  std::vector<uint32_t> meta(4 * window, 0), ranges, entries;
  for (uint32_t g = 0; g < window; ++g) {
    if (groups[g].empty()) continue;
    uint64_t bc = 1;
    while (bc < 2 * (uint64_t)groups[g].size()) bc <<= 1;
    const uint32_t boff = Out::fit(ranges.size() / 2, "prefilter bucket offset");
    std::vector<std::vector<uint32_t>> buckets(bc);
    for (uint32_t i : groups[g]) buckets[hash[i] & (bc - 1)].push_back(i);
    for (const auto& bk : buckets) {
      ranges.push_back(Out::fit(entries.size() / 2, "prefilter entry offset"));
      ranges.push_back((uint32_t)bk.size());
      for (uint32_t i : bk) {
        entries.push_back(hash[i]);
        entries.push_back(i);
      }
    }
    meta[4 * g + 0] = boff;
    meta[4 * g + 1] = Out::fit(bc - 1, "prefilter bucket count");
    meta[4 * g + 2] = (uint32_t)bc;
  }
  o.u32(window);
  o.raw(meta.data(), meta.size() * 4);
  o.u32((uint32_t)(ranges.size() / 2));
  o.raw(ranges.data(), ranges.size() * 4);
  o.u32((uint32_t)(entries.size() / 2));
  o.raw(entries.data(), entries.size() * 4);
}

// --------------------------------------------------------------------------
// Reader
// --------------------------------------------------------------------------

struct Cur {
  const uint8_t* d;
  size_t n;
  size_t off = 0;
  size_t remaining() const { return n - off; }
  const uint8_t* take(size_t k) {
    if (k > remaining()) corrupt("is truncated");
    const uint8_t* p = d + off;
    off += k;
    return p;
  }
  uint32_t u32() {
    const uint8_t* p = take(4);
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
  }
  // A declared element count must fit in what is left (mod.rs ensure_count_fits).
  void fits(uint64_t count, uint64_t size, const char* what) {
    if (count * size > remaining())
      corrupt(std::string(what) + " declares " + std::to_string(count * size) + " bytes but only " +
              std::to_string(remaining()) + " bytes remain");
  }
  std::string bytes() {
    const uint32_t len = u32();
    const uint8_t* p = take(len);
    return std::string(reinterpret_cast<const char*>(p), len);
  }
  std::vector<uint32_t> vec(const char* what) {
    const uint32_t c = u32();
    fits(c, 4, what);
    std::vector<uint32_t> v(c);
    if (c) std::memcpy(v.data(), take((size_t)c * 4), (size_t)c * 4);
    return v;
  }
  void cmap(std::array<uint8_t, 256>& m) { std::memcpy(m.data(), take(256), 256); }
};

// A version-5 regex section: the anchored DFA of one regex (states, classes,
// four start states, class map, table of states x (classes + 1)).
AnchoredDfa read_dfa(Cur& c) {
  AnchoredDfa d;
  d.nstates = c.u32();
  d.nclass = c.u32();
  for (auto& st : d.start) st = c.u32();
  c.cmap(d.cmap);
  d.table = c.vec("regex DFA table");
  if (d.nstates < 2 || d.nstates > DFA_STATE_MASK) corrupt("regex DFA state count is invalid");
  if (d.nclass < 1 || d.nclass > 256) corrupt("regex DFA class count is invalid");
  for (uint32_t st : d.start)
    if (st == 0 || st >= d.nstates) corrupt("regex DFA start state is out of range");
  for (int b = 0; b < 256; ++b)
    if (d.cmap[b] >= d.nclass) corrupt("regex DFA byte class points outside the class range");
  const uint64_t W = (uint64_t)d.nclass + 1;
  if ((uint64_t)d.nstates * W != d.table.size()) corrupt("regex DFA table size is inconsistent");
  for (uint32_t e : d.table)
    if ((e & DFA_STATE_MASK) >= d.nstates) corrupt("regex DFA transition points outside the state range");
  for (uint64_t s = 0; s < d.nstates; ++s)
    if (!(d.table[s * W + d.nclass] & DFA_DEAD)) corrupt("regex DFA end-of-text entry is not final");
  for (uint32_t st : d.start)
    for (uint64_t x = 0; x < W; ++x)
      if (d.table[st * W + x] & DFA_MATCH) d.can_empty = true;
  return d;
}

// A version-4 regex section (one start flag, no end-of-text column): its
// shape is checked and it is skipped; the set is rebuilt from its sources.
void skip_dfa_v4(Cur& c) {
  const uint32_t nstates = c.u32(), nclass = c.u32(), sm = c.u32();
  std::array<uint8_t, 256> cmap{};
  c.cmap(cmap);
  const std::vector<uint32_t> table = c.vec("regex DFA table");
  if (nstates < 2 || nclass < 1 || nclass > 256 || sm > 1 || (uint64_t)nstates * nclass != table.size())
    corrupt("regex DFA section is inconsistent");
}

// The version-4 factor automaton section.
AcTables read_ac(Cur& c) {
  AcTables t;
  const uint32_t ok = c.u32();
  t.states = c.u32();
  t.classes = c.u32();
  t.first_out = c.u32();
  c.cmap(t.cmap);
  t.table = c.vec("factor automaton table");
  t.out_start = c.vec("factor automaton output ranges");
  t.out_list = c.vec("factor automaton outputs");
  if (ok > 1) corrupt("factor automaton flag is invalid");
  t.ok = ok != 0;
  if (!t.ok) {
    if (t.states || t.classes || t.first_out || !t.table.empty() || !t.out_start.empty() ||
        !t.out_list.empty())
      corrupt("factor automaton section is inconsistent");
    return t;
  }
  if (t.states < 1) corrupt("factor automaton state count is invalid");
  if (t.classes < 1 || t.classes > 256) corrupt("factor automaton class count is invalid");
  if (t.first_out < 1 || t.first_out > t.states) corrupt("factor automaton output states are invalid");
  for (int b = 0; b < 256; ++b)
    if (t.cmap[b] >= t.classes) corrupt("factor automaton byte class points outside the class range");
  if ((uint64_t)t.states * t.classes != t.table.size()) corrupt("factor automaton table size is inconsistent");
  for (uint32_t e : t.table)
    if (e >= t.states) corrupt("factor automaton transition points outside the state range");
  if (t.out_start.size() != (size_t)(t.states - t.first_out) + 1 || t.out_start[0] != 0 ||
      t.out_start.back() != t.out_list.size())
    corrupt("factor automaton output ranges are inconsistent");
  for (size_t i = 1; i < t.out_start.size(); ++i)
    if (t.out_start[i] < t.out_start[i - 1]) corrupt("factor automaton output ranges are inconsistent");
  return t;
}

// A reference (version 1-3) regex section: regex-automata tables, read for
// their structure only (the pattern is rebuilt from its source).
void skip_reference_dfa(Cur& c) {
  (void)c.vec("u32 vector");  // transition_table
  (void)c.vec("u32 vector");  // match_list_pointers
  (void)c.vec("u32 vector");  // match_lists
  (void)c.vec("u32 vector");  // pattern_lengths
  (void)c.take(12);           // start_state, class_count, eoi_class
  (void)c.take(1024);         // byte_classes
  (void)c.bytes();            // native DFA bytes
  (void)c.vec("u32 vector");  // native_original_ids
}

}  // namespace

// Table-driven, eight bytes per step ("slicing-by-8", Kounavis and Berry,
// 2005): T[0] is the bytewise table of the reflected polynomial and
// T[k][i] = (T[k-1][i] >> 8) ^ T[0][T[k-1][i] & 0xFF] advances a byte k more
// positions.
uint32_t crc32_ieee(const uint8_t* p, size_t n) {
  static uint32_t T[8][256];
  static const bool init = [] {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k) c = (c & 1) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
      T[0][i] = c;
    }
    for (int k = 1; k < 8; ++k)
      for (uint32_t i = 0; i < 256; ++i) T[k][i] = (T[k - 1][i] >> 8) ^ T[0][T[k - 1][i] & 0xFF];
    return true;
  }();
  (void)init;
  uint32_t crc = 0xFFFFFFFFu;
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint32_t lo, hi;
    std::memcpy(&lo, p + i, 4);
    std::memcpy(&hi, p + i + 4, 4);
    lo ^= crc;
    crc = T[7][lo & 0xFF] ^ T[6][(lo >> 8) & 0xFF] ^ T[5][(lo >> 16) & 0xFF] ^ T[4][lo >> 24] ^
          T[3][hi & 0xFF] ^ T[2][(hi >> 8) & 0xFF] ^ T[1][(hi >> 16) & 0xFF] ^ T[0][hi >> 24];
  }
  for (; i < n; ++i) crc = T[0][(crc ^ p[i]) & 0xFF] ^ (crc >> 8);
  return ~crc;
}

std::vector<uint8_t> index_serialize(const PatternsHost& h, uint32_t version) {
  if (version != kIndexVersionRef && version != kIndexVersion)
    throw CompileError{WS_ERR_INVALID_ARGUMENT, "unsupported index version"};
  if (version == kIndexVersionRef && !h.rx.empty())
    throw CompileError{WS_ERR_INVALID_ARGUMENT,
                       "a version-3 index stores regex-automata DFAs, which this library does not "
                       "produce. Fix: write version 5, or keep the set to literals"};
  for (uint32_t i = 0; i < h.npat; ++i)
    if (h.has_name[i] && !valid_utf8(h.names[i]))
      throw CompileError{WS_ERR_INVALID_ARGUMENT, "pattern name " + std::to_string(i) + " is not valid UTF-8"};
  for (size_t r = 0; r < h.rx_src.size(); ++r)
    if (!valid_utf8(h.rx_src[r]))
      throw CompileError{WS_ERR_INVALID_ARGUMENT,
                         "regex " + std::to_string(h.rx_pid[r]) + " is not valid UTF-8 and cannot be stored"};
  Out o;
  // Header (builder.rs:23-33).
  o.raw(kMagic, 8);
  o.u32(version);
  o.u32(h.ci ? 1 : 0);
  o.u32(h.npat);
  o.u32(Out::fit(h.lits.size(), "literal count"));
  o.u32(Out::fit(h.rx.size(), "regex dfa count"));
  o.u32(kIndexHashWindow);
  // Packed literal bytes and their (start, len) spans (builder.rs:35-36).
  uint64_t total = 0;
  for (const auto& l : h.lits) total += l.size();
  o.u32(Out::fit(total, "payload length"));
  for (const auto& l : h.lits) o.raw(l.data(), l.size());
  o.u32((uint32_t)h.lits.size());
  uint64_t at = 0;
  for (const auto& l : h.lits) {
    o.u32((uint32_t)at);
    o.u32((uint32_t)l.size());
    at += l.size();
  }
  // Names: length-prefixed UTF-8, or u32::MAX for none (builder.rs:122-136).
  o.u32(h.npat);
  for (uint32_t i = 0; i < h.npat; ++i) {
    if (h.has_name[i]) o.bytes(h.names[i].data(), h.names[i].size(), "name length");
    else o.u32(kNoName);
  }
  // Regex sources with their pattern ids (builder.rs:139-149).
  o.u32((uint32_t)h.rx_src.size());
  for (size_t r = 0; r < h.rx_src.size(); ++r) {
    o.u32(h.rx_pid[r]);
    o.bytes(h.rx_src[r].data(), h.rx_src[r].size(), "payload length");
  }
  // Literal pattern ids, then the prefilter table (builder.rs:40-41).
  o.vec(h.lit_pid, "usize slice length");
  write_prefilter_table(o, h.lits, kIndexHashWindow);
  if (version == kIndexVersion) {
    for (const AnchoredDfa& d : h.rx) {
      o.u32(d.nstates);
      o.u32(d.nclass);
      for (uint32_t st : d.start) o.u32(st);
      o.raw(d.cmap.data(), 256);
      o.vec(d.table, "regex DFA table");
    }
    static const std::vector<uint32_t> none;
    o.u32(h.ac_ok ? 1 : 0);
    o.u32(h.ac_ok ? h.ac_states : 0);
    o.u32(h.ac_ok ? h.ac_classes : 0);
    o.u32(h.ac_ok ? h.ac_first_out : 0);
    std::array<uint8_t, 256> zero{};
    o.raw(h.ac_ok ? h.ac_cmap.data() : zero.data(), 256);
    o.vec(h.ac_ok ? h.ac_table : none, "factor automaton table");
    o.vec(h.ac_ok ? h.ac_out_start : none, "factor automaton output ranges");
    o.vec(h.ac_ok ? h.ac_out_list : none, "factor automaton outputs");
  }
  // CRC-32 trailer over everything before it (builder.rs:53-58).
  const uint32_t crc = crc32_ieee(o.b.data(), o.b.size());
  o.u32(crc);
  return std::move(o.b);
}

IndexContents index_parse(const uint8_t* data, size_t len) {
  IndexContents ic;
  IndexHeader& hd = ic.hdr;
  // Versions 3 and 4 end in a CRC-32 of everything before it (mod.rs:96-139).
  size_t end = len;
  if (len >= 16) {
    const uint32_t v = (uint32_t)data[8] | ((uint32_t)data[9] << 8) | ((uint32_t)data[10] << 16) |
                       ((uint32_t)data[11] << 24);
    if (v == kIndexVersionRef || v == kIndexVersionOld || v == kIndexVersion) {
      const uint32_t stored = (uint32_t)data[len - 4] | ((uint32_t)data[len - 3] << 8) |
                              ((uint32_t)data[len - 2] << 16) | ((uint32_t)data[len - 1] << 24);
      const uint32_t computed = crc32_ieee(data, len - 4);
      if (stored != computed) {
        char msg[96];
        std::snprintf(msg, sizeof msg, "integrity check failed (CRC mismatch: stored=%#010X, computed=%#010X)",
                      stored, computed);
        corrupt(msg);
      }
      end = len - 4;
    }
  }
  Cur c{data, end};
  if (std::memcmp(c.take(8), kMagic, 8) != 0) corrupt("header is invalid");
  hd.version = c.u32();
  if (hd.version > kIndexVersion) corrupt("version " + std::to_string(hd.version) + " is unsupported");
  if (hd.version == 0) corrupt("version is invalid");
  hd.ci = c.u32() != 0;
  hd.pattern_count = c.u32();
  hd.literal_count = c.u32();
  hd.regex_count = c.u32();
  hd.hash_window_len = c.u32();

  const std::string packed = c.bytes();
  const uint32_t noff = c.u32();
  c.fits(noff, 8, "pair vector");
  std::vector<std::pair<uint32_t, uint32_t>> offsets(noff);
  for (auto& o : offsets) {
    o.first = c.u32();
    o.second = c.u32();
  }
  if (noff != hd.literal_count) corrupt("literal metadata is inconsistent");

  const uint32_t nnames = c.u32();
  c.fits(nnames, 4, "name table");
  std::vector<std::string> names(nnames);
  std::vector<uint8_t> has_name(nnames, 0);
  for (uint32_t i = 0; i < nnames; ++i) {
    const uint32_t nl = c.u32();
    if (nl == kNoName) continue;
    names[i].assign(reinterpret_cast<const char*>(c.take(nl)), nl);
    if (!valid_utf8(names[i])) corrupt("name is not valid UTF-8");
    has_name[i] = 1;
  }
  if (nnames != hd.pattern_count) corrupt("name metadata is inconsistent");

  // Regex sources, version 2 on (mod.rs:183-189, 581-621).
  const bool have_sources = hd.version >= 2;
  std::vector<std::pair<uint32_t, std::string>> rx_src;
  if (have_sources) {
    const uint32_t nrx = c.u32();
    c.fits(nrx, 8, "regex pattern table");
    std::vector<uint8_t> seen(hd.pattern_count, 0);
    for (uint32_t r = 0; r < nrx; ++r) {
      const uint32_t pid = c.u32();
      if (pid >= hd.pattern_count) corrupt("regex pattern ID is out of bounds");
      if (seen[pid]) corrupt("contains duplicate regex pattern IDs");
      seen[pid] = 1;
      std::string src = c.bytes();
      if (!valid_utf8(src)) corrupt("regex pattern is not valid UTF-8");
      rx_src.emplace_back(pid, std::move(src));
    }
  }

  const std::vector<uint32_t> lit_ids = c.vec("u32 vector");
  if (lit_ids.size() != hd.literal_count) corrupt("literal ID table is inconsistent");

  // Prefilter table: structure and bucket ranges (mod.rs:623-679, 767-790).
  {
    const uint32_t nmeta = c.u32();
    c.fits(nmeta, 16, "prefix meta table");
    (void)c.take((size_t)nmeta * 16);
    const uint32_t nbuckets = c.u32();
    c.fits(nbuckets, 8, "bucket range table");
    const uint8_t* br = c.take((size_t)nbuckets * 8);
    const uint32_t nentries = c.u32();
    c.fits(nentries, 8, "prefilter entry table");
    (void)c.take((size_t)nentries * 8);
    for (uint32_t b = 0; b < nbuckets; ++b) {
      const uint8_t* q = br + 8 * b;
      const uint64_t st = (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16) | ((uint32_t)q[3] << 24);
      const uint64_t n = (uint32_t)q[4] | ((uint32_t)q[5] << 8) | ((uint32_t)q[6] << 16) | ((uint32_t)q[7] << 24);
      if (st + n > nentries)
        corrupt("prefilter bucket range #" + std::to_string(b) + " points outside the entry table");
    }
  }

  // Literal spans inside the packed bytes (mod.rs:204-219).
  for (const auto& o : offsets) {
    const uint64_t e = (uint64_t)o.first + o.second;
    if (e > 0xFFFFFFFFull) corrupt("literal offset overflows");
    if (e > packed.size()) corrupt("literal offset points outside the packed byte buffer");
  }

  // Regex sections.
  if (hd.version == kIndexVersion) {
    c.fits(hd.regex_count, 16, "regex DFA sections");
    ic.pre.rx.reserve(hd.regex_count);
    for (uint32_t r = 0; r < hd.regex_count; ++r) ic.pre.rx.push_back(read_dfa(c));
    ic.pre.ac = read_ac(c);
    hd.prebuilt = true;
  } else if (hd.version == kIndexVersionOld) {
    c.fits(hd.regex_count, 16, "regex DFA sections");
    for (uint32_t r = 0; r < hd.regex_count; ++r) skip_dfa_v4(c);
    (void)read_ac(c);
  } else {
    for (uint32_t r = 0; r < hd.regex_count; ++r) skip_reference_dfa(c);
  }
  if (c.remaining() != 0) corrupt("contains trailing bytes");

  // Rebuild the pattern list (to_pattern_set, mod.rs:287-372).
  if (!have_sources && hd.regex_count > 0)
    corrupt("lacks serialized regex patterns (version 1 stores only regex DFAs)");
  const uint32_t np = hd.pattern_count;
  std::vector<int64_t> rx_of(np, -1), lit_of(np, -1);
  for (size_t r = 0; r < rx_src.size(); ++r) rx_of[rx_src[r].first] = (int64_t)r;
  for (uint32_t l = 0; l < hd.literal_count; ++l) {
    const uint32_t pid = lit_ids[l];
    if (pid >= np) corrupt("literal pattern ID is out of bounds");
    if (lit_of[pid] >= 0) corrupt("contains duplicate literal IDs");
    lit_of[pid] = l;
  }
  ic.src.resize(np);
  for (uint32_t pid = 0; pid < np; ++pid) {
    PatternSrc& s = ic.src[pid];
    if (rx_of[pid] >= 0) {
      s.kind = WS_PATTERN_REGEX;
      s.bytes = std::move(rx_src[rx_of[pid]].second);
      s.has_name = has_name[pid] != 0;
      if (s.has_name) s.name = names[pid];
      continue;
    }
    if (lit_of[pid] < 0) corrupt("has missing pattern metadata");
    const auto& o = offsets[lit_of[pid]];
    s.kind = WS_PATTERN_LITERAL;
    s.bytes = packed.substr(o.first, o.second);
    // A named literal keeps its name only when its bytes are UTF-8
    // (to_pattern_set rebuilds other literals with literal_bytes).
    s.has_name = has_name[pid] && valid_utf8(s.bytes);
    if (s.has_name) s.name = names[pid];
  }
  return ic;
}

}  // namespace ws
