// Pattern-set compilation: literals, regex DFAs, prefilter factors, the
// rolling-hash tables and the dense factor automaton.
#include "patterns.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <deque>
#include <map>

#include "../../include/warpstate.h"

namespace ws {
namespace {

constexpr uint32_t kAcMaxLen = 12;        // factor DFA recognises literal prefixes up to this
constexpr uint32_t kAcMaxStates = 1u << 22;
constexpr size_t kNfaMaxStates = 200000;  // per regex
constexpr size_t kDfaMaxStates = 50000;   // per regex
constexpr int kFactorMaxLen = 8;
constexpr int kFactorFanout = 4;
constexpr int kFactorMaxCount = 256;

inline uint8_t fold(uint8_t b, bool ci) { return (ci && b >= 'A' && b <= 'Z') ? b + 32 : b; }

uint32_t ceil_log2(uint64_t x) {
  uint32_t r = 0;
  while ((1ull << r) < x) ++r;
  return r;
}

void build_hash(PatternsHost& ps) {
  const uint32_t F = (uint32_t)ps.fac.size();
  // Window lengths with a kernel build: every position for these...
  static const uint32_t kAllowed[] = {1, 2, 3, 4, 6, 8, 12, 16};
  // ...and every second position for these (a factor needs K + 1 bytes).
  static const uint32_t kAllowed2[] = {5, 7, 11, 15};
  uint32_t K = 1;
  for (uint32_t k : kAllowed) if (k <= ps.min_flen) K = k;
  // Half the windows at one byte less: half the filter work per byte, with
  // two grams per factor (factors of 6+ bytes) -- when the shorter grams
  // stay selective. Over text drawn from the factors' own A distinct bytes,
  // about 2F / A^K of the stride-2 windows hold a gram; above 1/2000 (DNA:
  // A = 4, 10,000 factors, K = 11: 1/210) the extra verification costs more
  // than the halved filter work saves.
  std::array<bool, 256> seen{};
  for (const Factor& f : ps.fac)
    for (char ch : f.bytes) seen[fold((uint8_t)ch, ps.ci)] = true;
  const double A = (double)std::max<size_t>(1, std::count(seen.begin(), seen.end(), true));
  for (uint32_t k : kAllowed2) {
    if (k + 1 > ps.min_flen || k + 1 < K) continue;
    if (2.0 * (double)F <= std::pow(A, (double)k) / 2000.0) K = k;
  }
  ps.K = K;
  ps.S = hash_stride(K);
  ps.B = K <= 2 ? 256u : 0x3779B1u;
  uint64_t bk = 1;
  for (uint32_t t = 0; t < K; ++t) bk = (bk * ps.B) & 0xFFFFFFu;
  ps.NBK = (uint32_t)((0x1000000u - bk) & 0xFFFFFFu);
  ps.dual = K >= 3;
  if (ps.dual) {
    // -B^(K-1) rolls the pair hash (every position: K - 1 pair terms),
    // -B^K the byte hash (every second position). The multipliers were
    // picked by simulating the filter's hit rate against that of a random
    // hash in the same layout: 10,000 DNA 12-mers over ACGT text 0.00300
    // hits per byte (random hash 0.00301, the pair 0xA3B1 / 0xC6A5
    // 0.00341), and about the random hash's rate for [a-z_] and printable
    // literal sets;
    // some odd multipliers mix the low bits so poorly that hits rise
    // fourfold.
    const uint32_t Bl = 0xD5DDu, Bh = 0xE70Bu;
    uint32_t pl = 1, ph = 1;
    for (uint32_t t = 0; t < (hash_stride(K) == 1 ? K - 1 : K); ++t) {
      pl = (pl * Bl) & 0xFFFFu;
      ph = (ph * Bh) & 0xFFFFu;
    }
    ps.BB = Bl | (Bh << 16);
    ps.BB2 = ((Bl * Bl) & 0xFFFFu) | (((Bh * Bh) & 0xFFFFu) << 16);
    ps.NBKK = ((0x10000u - pl) & 0xFFFFu) | (((0x10000u - ph) & 0xFFFFu) << 16);
  }
  const uint32_t S = ps.S, E = F * S;  // verification entries: S grams per factor
  if (K <= 2) {
    // Exact filter over the raw 1- or 2-byte key: one bit per key value.
    ps.wbits = 8 * K - 5;
    ps.wshift = 5;
    ps.s2 = 0;
  } else {
    // Blocked bloom filter over the 32-bit dual hash D: the word is
    // D[wbits+1:2], its two bits D[20:16] and D[28:24]; about 256 bits per
    // gram, 1 KiB .. 128 KiB.
    uint32_t lb = std::min(20u, std::max(13u, ceil_log2(E) + 8));
    ps.wbits = lb - 5;
    ps.wshift = 2;
    ps.s2 = 5;
  }
  ps.bloom.assign(1u << ps.wbits, 0);
  // Verification buckets: eight per entry while the table (64 bytes a
  // bucket) stays within 64 MiB, at least two per entry. The sparser table
  // sends fewer filter hits to a second slot or a crowded bucket (measured
  // against two per entry: 10,000 DNA 12-mers 0.2741 -> 0.2709 ms at four,
  // 200 regexes' chunks 0.4238 -> 0.4178 ms at eight).
  ps.nbits = std::max({1u, std::min(ceil_log2(8ull * E), 20u), ceil_log2(2ull * E)});
  const uint32_t nb = 1u << ps.nbits;
  ps.ent_fac.resize(E);
  ps.ent_off.resize(E);
  ps.ent_hash.resize(E);
  std::vector<uint32_t> bucket(E), cnt(nb + 1, 0);
  for (uint32_t f = 0; f < F; ++f) {
    for (uint32_t o = 0; o < S; ++o) {
      const uint32_t e = f * S + o;
      const uint8_t* g = (const uint8_t*)ps.fac[f].bytes.data() + o;
      const uint32_t h = rolling_hash(g, K, ps.B, ps.ci);
      const uint32_t key = window_key(g, K, ps.ci);
      ps.ent_fac[e] = f;
      ps.ent_off[e] = o;
      ps.ent_hash[e] = key;
      if (ps.dual) {
        const uint32_t d = S == 1 ? pair_hash(g, K, ps.BB, ps.ci) : dual_hash(g, K, ps.BB, ps.ci);
        const uint32_t w = (d >> 2) & ((1u << ps.wbits) - 1);
        ps.bloom[w] |= 1u << ((d >> 16) & 31);
        ps.bloom[w] |= 1u << ((d >> 24) & 31);
      } else {
        const uint32_t w = (h >> ps.wshift) & ((1u << ps.wbits) - 1);
        ps.bloom[w] |= 1u << (h & 31);
        ps.bloom[w] |= 1u << ((h >> ps.s2) & 31);
      }
      bucket[e] = (uint32_t)((key * 0x9E3779B1u) >> (32 - ps.nbits));
      cnt[bucket[e] + 1]++;
    }
  }
  for (uint32_t b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
  ps.bstart = cnt;
  ps.bent.resize(E);
  std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
  for (uint32_t e = 0; e < E; ++e) ps.bent[fill[bucket[e]]++] = e;
  ps.hash_ok = true;
}

void build_ac(PatternsHost& ps) {
  // Trie over the factor strings (folded when case-insensitive).
  std::vector<std::map<uint8_t, uint32_t>> kids(1);
  std::vector<std::vector<uint32_t>> own(1);
  std::vector<uint8_t> used(256, 0);
  ps.ac_maxlen = 0;
  for (uint32_t f = 0; f < ps.fac.size(); ++f) {
    const Factor& fc = ps.fac[f];
    uint32_t node = 0;
    for (uint32_t t = 0; t < fc.ac_len; ++t) {
      uint8_t b = fold((uint8_t)fc.bytes[t], ps.ci);
      used[b] = 1;
      auto it = kids[node].find(b);
      if (it == kids[node].end()) {
        if (kids.size() >= kAcMaxStates) { ps.ac_ok = false; return; }
        uint32_t nn = (uint32_t)kids.size();
        kids.emplace_back();
        own.emplace_back();
        kids[node][b] = nn;
        node = nn;
      } else {
        node = it->second;
      }
    }
    own[node].push_back(f);
    ps.ac_maxlen = std::max(ps.ac_maxlen, fc.ac_len);
  }
  const uint32_t S = (uint32_t)kids.size();
  // Byte classes: every used byte is its own class; the rest share class 0.
  uint32_t C = 1;
  std::array<uint8_t, 256> cmap{};
  std::vector<int> rep(1, -1);
  for (int b = 0; b < 256; ++b)
    if (used[b]) { cmap[b] = (uint8_t)C++; rep.push_back(b); }
  if (C > 256) { ps.ac_ok = false; return; }
  if (ps.ci)
    for (int b = 'A'; b <= 'Z'; ++b) cmap[b] = cmap[b + 32];
  // BFS: failure links, dense goto, inherited outputs.
  std::vector<uint32_t> fail(S, 0), order;
  std::vector<uint32_t> delta((size_t)S * C, 0);
  std::vector<std::vector<uint32_t>> out(S);
  std::deque<uint32_t> q;
  order.push_back(0);
  out[0] = own[0];
  for (uint32_t c = 1; c < C; ++c) {
    auto it = kids[0].find((uint8_t)rep[c]);
    if (it != kids[0].end()) {
      delta[c] = it->second;
      fail[it->second] = 0;
      q.push_back(it->second);
    }
  }
  while (!q.empty()) {
    uint32_t u = q.front();
    q.pop_front();
    order.push_back(u);
    out[u] = own[u];
    for (uint32_t f : out[fail[u]]) out[u].push_back(f);
    for (uint32_t c = 0; c < C; ++c) {
      uint32_t via_fail = delta[(size_t)fail[u] * C + c];
      if (c == 0) { delta[(size_t)u * C] = via_fail; continue; }
      auto it = kids[u].find((uint8_t)rep[c]);
      if (it != kids[u].end()) {
        uint32_t v = it->second;
        fail[v] = via_fail;
        delta[(size_t)u * C + c] = v;
        q.push_back(v);
      } else {
        delta[(size_t)u * C + c] = via_fail;
      }
    }
  }
  // Renumber: states with outputs last so the kernel tests `s >= first_out`.
  std::vector<uint32_t> newid(S);
  uint32_t next = 0;
  for (uint32_t u : order) if (out[u].empty()) newid[u] = next++;
  ps.ac_first_out = next;
  for (uint32_t u : order) if (!out[u].empty()) newid[u] = next++;
  ps.ac_states = S;
  ps.ac_classes = C;
  ps.ac_cmap = cmap;
  ps.ac_table.assign((size_t)S * C, 0);
  for (uint32_t u = 0; u < S; ++u)
    for (uint32_t c = 0; c < C; ++c)
      ps.ac_table[(size_t)newid[u] * C + c] = newid[delta[(size_t)u * C + c]];
  std::vector<uint32_t> by_new(S);
  for (uint32_t u = 0; u < S; ++u) by_new[newid[u]] = u;
  ps.ac_out_start.assign(S - ps.ac_first_out + 1, 0);
  ps.ac_out_list.clear();
  for (uint32_t s = ps.ac_first_out; s < S; ++s) {
    ps.ac_out_start[s - ps.ac_first_out] = (uint32_t)ps.ac_out_list.size();
    auto& o = out[by_new[s]];
    std::vector<uint32_t> sorted(o);
    std::sort(sorted.begin(), sorted.end());
    for (uint32_t f : sorted) ps.ac_out_list.push_back(f);
  }
  ps.ac_out_start[S - ps.ac_first_out] = (uint32_t)ps.ac_out_list.size();
  ps.ac_ok = newid[0] == 0;  // the root must be state 0 and carry no output
}

// Takes a stored factor automaton. The reader has checked its shape (table
// entries are states, classes cover the byte map, output ranges are ordered);
// here its outputs are checked against this set's factors.
void install_ac(PatternsHost& ps, AcTables& t) {
  ps.ac_ok = false;
  ps.ac_maxlen = 0;
  for (const Factor& f : ps.fac) ps.ac_maxlen = std::max(ps.ac_maxlen, f.ac_len);
  if (!t.ok) return;
  const uint32_t F = (uint32_t)ps.fac.size();
  for (uint32_t f : t.out_list)
    if (f >= F)
      throw CompileError{WS_ERR_PATTERN_COMPILATION_FAILED,
                         "compiled index factor automaton names a factor out of bounds"};
  ps.ac_states = t.states;
  ps.ac_classes = t.classes;
  ps.ac_first_out = t.first_out;
  ps.ac_cmap = t.cmap;
  ps.ac_table = std::move(t.table);
  ps.ac_out_start = std::move(t.out_start);
  ps.ac_out_list = std::move(t.out_list);
  ps.ac_ok = true;
}

}  // namespace

uint32_t dual_hash(const uint8_t* p, uint32_t K, uint32_t BB, bool ci) {
  const uint32_t Bl = BB & 0xFFFFu, Bh = BB >> 16;
  uint32_t hl = 0, hh = 0;
  for (uint32_t t = 0; t < K; ++t) {
    const uint32_t b = (ci && p[t] >= 'A' && p[t] <= 'Z') ? p[t] + 32u : p[t];
    hl = (hl * Bl + b) & 0xFFFFu;
    hh = (hh * Bh + b) & 0xFFFFu;
  }
  return hl | (hh << 16);
}

uint32_t pair_hash(const uint8_t* p, uint32_t K, uint32_t BB, bool ci) {
  const uint32_t Bl = BB & 0xFFFFu, Bh = BB >> 16;
  uint32_t hl = 0, hh = 0;
  for (uint32_t t = 0; t + 1 < K; ++t) {
    const uint32_t x = fold(p[t], ci) | ((uint32_t)fold(p[t + 1], ci) << 8);
    hl = (hl * Bl + x) & 0xFFFFu;
    hh = (hh * Bh + x) & 0xFFFFu;
  }
  return hl | (hh << 16);
}

uint32_t window_key(const uint8_t* p, uint32_t K, bool ci) {
  uint32_t h = 0;
  for (uint32_t j = 0; 3 * j < K; ++j) {
    uint32_t c = 0;
    for (uint32_t i = 0; i < 3 && 3 * j + i < K; ++i) c |= (uint32_t)fold(p[3 * j + i], ci) << (8 * i);
    h = j == 0 ? c : (h & 0xFFFFFFu) * kWindowKeyMul + c;
  }
  return h & 0x7FFFFFFFu;
}

uint32_t rolling_hash(const uint8_t* p, uint32_t K, uint32_t B, bool ci) {
  uint32_t h = 0;
  for (uint32_t t = 0; t < K; ++t) h = (h * B + fold(p[t], ci)) & 0xFFFFFFu;
  return h;
}

PatternsHost compile_patterns(const std::vector<PatternSrc>& src, bool case_insensitive,
                              Prebuilt* pre) {
  if (src.empty()) throw CompileError{WS_ERR_EMPTY_PATTERN_SET, "pattern set is empty"};
  if (src.size() > 0x7FFFFFFFull)
    throw CompileError{WS_ERR_PATTERN_SET_TOO_LARGE, "too many patterns"};
  PatternsHost ps;
  ps.npat = (uint32_t)src.size();
  ps.ci = case_insensitive;
  ps.names.resize(ps.npat);
  ps.has_name.assign(ps.npat, 0);
  for (uint32_t i = 0; i < ps.npat; ++i)
    if (src[i].has_name) {
      ps.names[i] = src[i].name;
      ps.has_name[i] = 1;
    }
  bool unbounded = false;
  uint64_t max_len = 0;
  for (uint32_t i = 0; i < ps.npat; ++i) {
    const PatternSrc& p = src[i];
    if (p.bytes.empty())
      throw CompileError{WS_ERR_EMPTY_PATTERN, "pattern " + std::to_string(i) + " is empty"};
    if (p.bytes.size() > 0xFFFFFFFFull)
      throw CompileError{WS_ERR_PATTERN_TOO_LARGE, "pattern " + std::to_string(i) + " too large"};
    if (p.kind == WS_PATTERN_LITERAL) {
      ps.lits.push_back(p.bytes);
      ps.lit_pid.push_back(i);
      max_len = std::max<uint64_t>(max_len, p.bytes.size());
      continue;
    }
    if (p.kind != WS_PATTERN_REGEX)
      throw CompileError{WS_ERR_INVALID_ARGUMENT, "unknown pattern kind"};
    Ast ast;
    try {
      ast = parse_regex(p.bytes, false);
    } catch (const ParseError& e) {
      throw CompileError{WS_ERR_PATTERN_COMPILATION_FAILED,
                         "invalid regex at index " + std::to_string(i) + ": " + e.msg};
    }
    if (is_pathological(ast))
      throw CompileError{WS_ERR_PATHOLOGICAL_REGEX,
                         "regex pattern " + std::to_string(i) + " contains nested repetitions"};
    std::string lit;
    if (extract_literal(ast, &lit) && !lit.empty()) {
      ps.lits.push_back(lit);
      ps.lit_pid.push_back(i);
      max_len = std::max<uint64_t>(max_len, lit.size());
      continue;
    }
    long mn, mx;
    length_bounds(ast, &mn, &mx);
    try {
      Nfa nfa = build_nfa(ast, kNfaMaxStates);
      const uint32_t r = (uint32_t)ps.rx.size();
      AnchoredDfa dfa;
      if (pre) {
        if (r >= pre->rx.size())
          throw CompileError{WS_ERR_PATTERN_COMPILATION_FAILED,
                             "compiled index has fewer regex DFA sections than regex patterns"};
        dfa = std::move(pre->rx[r]);
      } else {
        dfa = build_anchored_dfa(nfa, kDfaMaxStates);
      }
      std::vector<std::string> fs = extract_factors(nfa, kFactorMaxLen, kFactorFanout, kFactorMaxCount);
      // An empty factor: a match may start with (almost) any byte, or be
      // empty. Such a regex takes every position as a candidate start.
      const bool scan_all = std::any_of(fs.begin(), fs.end(), [](const std::string& f) { return f.empty(); });
      ps.rx.push_back(std::move(dfa));
      ps.rx_src.push_back(p.bytes);
      ps.rx_pid.push_back(i);
      ps.rx_maxlen.push_back(mx);
      if (mx < 0) unbounded = true;
      else max_len = std::max<uint64_t>(max_len, (uint64_t)mx);
      if (scan_all)
        ps.rx_scan.push_back(r);
      else
        for (auto& s : fs) ps.fac.push_back({s, 1u, r, (uint32_t)s.size()});
    } catch (const ParseError& e) {
      throw CompileError{WS_ERR_PATTERN_COMPILATION_FAILED,
                         "regex " + std::to_string(i) + ": " + e.msg};
    }
  }
  // Literal factors first (verification string = the whole literal).
  std::vector<Factor> fac;
  for (uint32_t l = 0; l < ps.lits.size(); ++l)
    fac.push_back({ps.lits[l], 0u, l, std::min<uint32_t>((uint32_t)ps.lits[l].size(), kAcMaxLen)});
  for (auto& f : ps.fac) fac.push_back(f);
  ps.fac = std::move(fac);
  ps.min_flen = ps.fac.empty() ? 0u : 0xFFFFFFFFu;
  ps.max_flen = 0;
  for (auto& f : ps.fac) {
    ps.min_flen = std::min<uint32_t>(ps.min_flen, (uint32_t)f.bytes.size());
    ps.max_flen = std::max<uint32_t>(ps.max_flen, (uint32_t)f.bytes.size());
  }
  ps.min_lit_len = 0;
  for (auto& l : ps.lits)
    ps.min_lit_len = ps.min_lit_len == 0 ? (uint32_t)l.size() : std::min<uint32_t>(ps.min_lit_len, (uint32_t)l.size());
  ps.max_match_len = unbounded ? 0xFFFFFFFFu : (uint32_t)std::min<uint64_t>(max_len, 0xFFFFFFFEull);

  const size_t nl = ps.lits.size(), nr = ps.rx.size();
  if (nl == 0) ps.strategy = STRAT_NONE;
  else if (ps.ci || nr > 0) ps.strategy = STRAT_LEFTMOST_LONGEST;
  else if (nl >= 2 && nl <= 8 &&
           std::all_of(ps.lits.begin(), ps.lits.end(), [](const std::string& s) { return s.size() <= 16; }))
    ps.strategy = STRAT_MULTI_MEMCHR;
  else if (nl > 5000) ps.strategy = STRAT_HASH_SHORTEST;
  else ps.strategy = STRAT_LEFTMOST_LONGEST;

  build_hash(ps);
  if (!pre) {
    build_ac(ps);
    return ps;
  }
  if (pre->rx.size() != ps.rx.size())
    throw CompileError{WS_ERR_PATTERN_COMPILATION_FAILED,
                       "compiled index has more regex DFA sections than regex patterns"};
  install_ac(ps, pre->ac);
  return ps;
}

}  // namespace ws
