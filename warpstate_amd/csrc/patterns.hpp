// Host-side compiled pattern set: what PatternSetBuilder::build produces in
// the reference (core/src/pattern/compiler.rs:97-340), laid out for the GPU.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "regex.hpp"

namespace ws {

// Literal reduction used by WS_MODE_STANDARD, as ScanStrategy::select picks it
// (core/src/specialize.rs:101-179, core/src/cpu/mod.rs:225-230).
enum Strategy : uint32_t {
  STRAT_NONE = 0,        // no literals
  STRAT_LEFTMOST_LONGEST = 1,  // Aho-Corasick LeftmostLongest (also single literal)
  STRAT_MULTI_MEMCHR = 2,      // 2..8 literals <= 16 bytes, no regex, case-sensitive
  STRAT_HASH_SHORTEST = 3,     // > 5000 literals, no regex, case-sensitive
};

// A prefilter factor: every match of its owner starts with `bytes`.
// Literal factors are the whole literal; regex factors come from
// extract_factors(). `ac_len` is the length the factor DFA recognises
// (literals are truncated to keep the automaton small).
struct Factor {
  std::string bytes;
  uint32_t kind;   // 0 = literal (owner = literal index), 1 = regex (owner = regex index)
  uint32_t owner;
  uint32_t ac_len;
};

struct PatternsHost {
  uint32_t npat = 0;
  bool ci = false;
  // Per pattern id: the optional name (PatternSetBuilder::named_*).
  std::vector<std::string> names;
  std::vector<uint8_t> has_name;
  std::vector<std::string> lits;
  std::vector<uint32_t> lit_pid;
  std::vector<uint32_t> rx_pid;
  std::vector<std::string> rx_src;  // regex source text, per regex
  std::vector<AnchoredDfa> rx;
  std::vector<long> rx_maxlen;  // -1 = unbounded
  // Regexes without a selective prefix (their first byte admits too many
  // values, or a match can be empty): every position of the input is a
  // candidate start, walked by k_rx_scan instead of coming from the
  // prefilter. Regex indices, ascending.
  std::vector<uint32_t> rx_scan;
  std::vector<Factor> fac;
  uint32_t min_flen = 0, max_flen = 0, min_lit_len = 0;
  uint32_t max_match_len = 0;  // 0xFFFFFFFF when some regex is unbounded
  uint32_t strategy = STRAT_NONE;

  // Rolling-hash prefilter over K-byte windows at every S-th position
  // (hash_stride(K)); with S = 2 every factor has at least K + 1 bytes and
  // both its K-grams at offsets 0 and 1 are in the filter.
  bool hash_ok = false;
  uint32_t K = 0, S = 1, B = 0, NBK = 0, wshift = 0, wbits = 0, nbits = 0, s2 = 0;
  // Bloom hash for K >= 3: two 16-bit polynomial hashes side by side
  // (multipliers packed low | high << 16, their squares, and -multiplier^K
  // likewise).
  bool dual = false;
  uint32_t BB = 0, BB2 = 0, NBKK = 0;
  std::vector<uint32_t> bloom;
  // Verification entries (factor, gram offset < S, 24-bit rolling hash of the
  // gram), bucketed by hash: bucket b holds bent[bstart[b] .. bstart[b+1]).
  std::vector<uint32_t> ent_fac, ent_off, ent_hash, bstart, bent;

  // Factor DFA (dense Aho-Corasick automaton over byte classes).
  bool ac_ok = false;
  uint32_t ac_states = 0, ac_classes = 0, ac_first_out = 0, ac_maxlen = 0;
  std::array<uint8_t, 256> ac_cmap{};
  std::vector<uint32_t> ac_table, ac_out_start, ac_out_list;
};

struct CompileError {
  int status;  // ws_status
  std::string msg;
};

struct PatternSrc {
  uint32_t kind;
  std::string bytes;
  std::string name;
  bool has_name = false;
};

// The factor automaton as stored in a compiled index.
struct AcTables {
  bool ok = false;
  uint32_t states = 0, classes = 0, first_out = 0;
  std::array<uint8_t, 256> cmap{};
  std::vector<uint32_t> table, out_start, out_list;
};

// Tables read from a compiled index in place of the two expensive steps of
// compilation: the subset construction of each regex DFA (one per regex, in
// pattern order) and the factor automaton. Their structure has been checked
// by the reader; compile_patterns checks them against the factor set.
struct Prebuilt {
  std::vector<AnchoredDfa> rx;
  AcTables ac;
};

// Throws CompileError. With `pre`, its tables are moved into the result
// instead of being built.
PatternsHost compile_patterns(const std::vector<PatternSrc>& src, bool case_insensitive,
                              Prebuilt* pre = nullptr);

// Window stride of the rolling-hash filter for window length K: the odd K
// (5, 7, 11, 15) test every second position, the others every position.
// Shared by the set compiler (which picks K) and the filter kernel.
constexpr uint32_t hash_stride(uint32_t K) { return (K >= 5 && (K & 1u)) ? 2u : 1u; }

// Rolling hash of the first K bytes (mod 2^24): the exact filter's key for
// K <= 2 (B = 256).
uint32_t rolling_hash(const uint8_t* p, uint32_t K, uint32_t B, bool ci);
// Verification-table key of a K-byte window (K <= 16), shared by host and
// kernel: the window's 3-byte little-endian chunks c_j (bytes past K zero),
// H = c_0, then H = (H mod 2^24) * kWindowKeyMul + c_j (mod 2^32) for each
// next chunk, top bit cleared (the table's empty and overflow markers have
// it set). On the device a chunk is one byte permute and a step one
// v_mad_u32_u24: 6 operations for 12 bytes.
constexpr uint32_t kWindowKeyMul = 0x9E3779u;
uint32_t window_key(const uint8_t* p, uint32_t K, bool ci);
// The 32-bit bloom hash of a K-byte window at every second position: low
// half sum b[t] Bl^(K-1-t), high half with Bh, each mod 2^16
// (BB = Bl | Bh << 16).
uint32_t dual_hash(const uint8_t* p, uint32_t K, uint32_t BB, bool ci);
// The same over the K - 1 little-endian byte pairs x[t] = b[t] | b[t+1] << 8
// (windows at every position): sum x[t] B^(K-2-t) per half.
uint32_t pair_hash(const uint8_t* p, uint32_t K, uint32_t BB, bool ci);

}  // namespace ws
