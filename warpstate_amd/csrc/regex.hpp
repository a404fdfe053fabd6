// Byte-regex front end for the GPU scan path.
//
// Semantics follow the reference's regex configuration: `regex::bytes` with
// `multi_line(true)` and `unicode(false)` (core/src/regex_engine.rs:7-17), so
// `.` is any byte except '\n', classes are ASCII, and matches are
// leftmost-first, greedy, non-overlapping per pattern
// (core/src/dfa/scan.rs:433-457).
//
// Zero-width assertions: `^` and `$` at line starts and ends (multi-line),
// `\A` and `\z` at the text's, `\b` and `\B` on the ASCII word class; and
// regexes that can match the empty string, iterated as the regex crate's
// find_iter does.
//
// Pipeline: parse -> AST -> Thompson NFA with priority-ordered splits ->
//   * an anchored leftmost-first DFA (ordered-subset construction), walked on
//     the device from each candidate start to find the match end;
//   * a prefix-free "factor" set: byte strings such that every match starts
//     with one of them (used by the literal-style prefilters).
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace ws {

struct ByteSet {
  std::array<uint64_t, 4> w{};
  void set(int b) { w[b >> 6] |= 1ull << (b & 63); }
  bool has(int b) const { return (w[b >> 6] >> (b & 63)) & 1; }
  void set_range(int lo, int hi) { for (int b = lo; b <= hi; ++b) set(b); }
  void merge(const ByteSet& o) { for (int i = 0; i < 4; ++i) w[i] |= o.w[i]; }
  void negate() { for (auto& x : w) x = ~x; }
  int count() const { int c = 0; for (auto x : w) c += __builtin_popcountll(x); return c; }
  int single() const {  // the only member, or -1
    if (count() != 1) return -1;
    for (int b = 0; b < 256; ++b) if (has(b)) return b;
    return -1;
  }
  bool operator==(const ByteSet& o) const { return w == o.w; }
};

// Assertion kinds (Ast::LOOK, Nfa::LOOK).
enum Look : int { LOOK_BOL = 0, LOOK_EOL, LOOK_BOT, LOOK_EOT, LOOK_WORDB, LOOK_NWORDB };

struct Ast {
  enum Kind { EMPTY, BYTES, CONCAT, ALT, REPEAT, GROUP, LOOK } kind = EMPTY;
  int look = 0;                // LOOK
  ByteSet set;                 // BYTES
  std::vector<Ast> kids;       // CONCAT / ALT / REPEAT(1) / GROUP(1)
  int rmin = 0, rmax = -1;     // REPEAT; rmax == -1 means unbounded
  bool greedy = true;
};

struct ParseError { std::string msg; };

// Parses `pattern`; throws ParseError. Unsupported constructs (look-around,
// backreferences, class set operations, verbose mode, Unicode classes,
// repeated assertions) are reported as errors.
Ast parse_regex(const std::string& pattern, bool case_insensitive);

// Nested repetition check of core/src/pattern/compiler.rs:381-405.
bool is_pathological(const Ast& a);
// Pure-literal check of core/src/pattern/compiler.rs:360-375.
bool extract_literal(const Ast& a, std::string* out);
// Minimum / maximum match length (max = -1 when unbounded).
void length_bounds(const Ast& a, long* mn, long* mx);

struct Nfa {
  enum Type : uint8_t { BYTES, SPLIT, EPS, MATCH, LOOK };
  // BYTES: `set` indexes sets; LOOK: `set` is the Look kind.
  struct State { Type t; int out = -1, out1 = -1; int set = -1; };
  std::vector<State> st;
  std::vector<ByteSet> sets;
  int start = -1;
};

// Builds the NFA; throws ParseError when it would exceed `max_states`.
Nfa build_nfa(const Ast& a, size_t max_states);

// Dense anchored leftmost-first DFA over byte classes, with the match
// delayed by one byte (a state knows the byte before it; whether a match ends
// at a position can depend on the byte after it). State 0 is dead. Rows have
// nclass + 1 entries: column c < nclass is the transition on a byte of class
// c, column nclass the end of the text. Entry = next | DFA_MATCH when a match
// ends at the current position (before that byte) | DFA_DEAD when next is 0
// (the end-of-text column is always DFA_DEAD).
// A walk from position p starts in start[ctx], ctx being what precedes p:
// 0 the text start, 1 '\n', 2 a word byte [0-9A-Za-z_], 3 any other byte.
struct AnchoredDfa {
  uint32_t nstates = 0, nclass = 0;
  std::array<uint8_t, 256> cmap{};
  std::vector<uint32_t> table;  // nstates * (nclass + 1)
  std::array<uint32_t, 4> start{};
  bool can_empty = false;       // some start state matches before any byte (an empty match)
};
constexpr uint32_t DFA_MATCH = 0x80000000u;
constexpr uint32_t DFA_DEAD = 0x40000000u;
constexpr uint32_t DFA_STATE_MASK = 0x3FFFFFFFu;
// The context a byte leaves for the position after it (see AnchoredDfa::start).
inline uint32_t dfa_ctx_of(int b) {
  if (b == '\n') return 1;
  const bool w = (b >= '0' && b <= '9') || (b >= 'A' && b <= 'Z') || (b >= 'a' && b <= 'z') || b == '_';
  return w ? 2 : 3;
}

AnchoredDfa build_anchored_dfa(const Nfa& n, size_t max_states);

// Whether the NFA has assertions.
bool has_look(const Nfa& n);

// Prefix-free factor strings (see file comment). Assertions are taken as
// satisfiable (the walk checks them). Strings are extended while
// the next-byte fanout stays <= max_fanout, up to max_len bytes and
// max_count strings in total.
std::vector<std::string> extract_factors(const Nfa& n, int max_len, int max_fanout,
                                         int max_count);

}  // namespace ws
