// Byte-regex front end: parser, NFA, anchored leftmost-first DFA, factors.
#include "regex.hpp"

#include <algorithm>
#include <map>
#include <unordered_map>

namespace ws {
namespace {

ByteSet range_set(int lo, int hi) { ByteSet s; s.set_range(lo, hi); return s; }

ByteSet digit_set() { return range_set('0', '9'); }
ByteSet word_set() {
  ByteSet s = range_set('0', '9');
  s.merge(range_set('A', 'Z'));
  s.merge(range_set('a', 'z'));
  s.set('_');
  return s;
}
ByteSet space_set() {  // ASCII whitespace: \t \n \v \f \r and space
  ByteSet s = range_set(9, 13);
  s.set(' ');
  return s;
}

void fold_case(ByteSet& s) {
  for (int b = 'a'; b <= 'z'; ++b) {
    int u = b - 32;
    if (s.has(b) || s.has(u)) { s.set(b); s.set(u); }
  }
}

Ast bytes_node(const ByteSet& s) { Ast a; a.kind = Ast::BYTES; a.set = s; return a; }

struct Flags { bool ci = false, dotall = false, swap_greed = false; };

struct Parser {
  const std::string& p;
  size_t i = 0;
  Flags f;

  explicit Parser(const std::string& pat) : p(pat) {}

  [[noreturn]] void fail(const std::string& m) {
    throw ParseError{m + " at offset " + std::to_string(i)};
  }
  bool eof() const { return i >= p.size(); }
  int peek(size_t k = 0) const { return i + k < p.size() ? (unsigned char)p[i + k] : -1; }

  Ast parse_alt() {
    std::vector<Ast> branches;
    branches.push_back(parse_concat());
    while (peek() == '|') { ++i; branches.push_back(parse_concat()); }
    if (branches.size() == 1) return std::move(branches[0]);
    Ast a; a.kind = Ast::ALT; a.kids = std::move(branches);
    return a;
  }

  Ast parse_concat() {
    std::vector<Ast> items;
    while (!eof() && peek() != '|' && peek() != ')') {
      bool flag_only = false;
      Ast atom = parse_atom(&flag_only);
      if (flag_only) continue;
      if (atom.kind == Ast::LOOK) {
        const int q = peek();
        if (q == '*' || q == '+' || q == '?' || q == '{') fail("repetition of an assertion");
        items.push_back(std::move(atom));
        continue;
      }
      items.push_back(parse_quant(std::move(atom)));
    }
    if (items.empty()) return Ast{};
    if (items.size() == 1) return std::move(items[0]);
    Ast a; a.kind = Ast::CONCAT; a.kids = std::move(items);
    return a;
  }

  bool parse_uint(int* out) {
    size_t st = i;
    long v = 0;
    while (peek() >= '0' && peek() <= '9') {
      v = v * 10 + (peek() - '0');
      if (v > 100000) fail("repetition count too large");
      ++i;
    }
    *out = (int)v;
    return i > st;
  }

  Ast parse_quant(Ast atom) {
    int c = peek();
    int mn, mx;
    if (c == '*') { mn = 0; mx = -1; ++i; }
    else if (c == '+') { mn = 1; mx = -1; ++i; }
    else if (c == '?') { mn = 0; mx = 1; ++i; }
    else if (c == '{') {
      ++i;
      if (!parse_uint(&mn)) fail("invalid counted repetition");
      mx = mn;
      if (peek() == ',') {
        ++i;
        if (peek() == '}') mx = -1;
        else if (!parse_uint(&mx)) fail("invalid counted repetition");
      }
      if (peek() != '}') fail("unclosed counted repetition");
      ++i;
      if (mx != -1 && mx < mn) fail("invalid repetition range");
      if (mn > 1000 || mx > 1000) fail("repetition count exceeds 1000");
    } else {
      return atom;
    }
    bool greedy = true;
    if (peek() == '?') { greedy = false; ++i; }
    if (f.swap_greed) greedy = !greedy;
    int n = peek();
    if (n == '*' || n == '+' || n == '?' || n == '{') fail("stacked repetition operator");
    if (mn == 1 && mx == 1) return atom;
    if (mx == 0) return Ast{};
    Ast r; r.kind = Ast::REPEAT; r.rmin = mn; r.rmax = mx; r.greedy = greedy;
    r.kids.push_back(std::move(atom));
    return r;
  }

  Ast literal_byte(int b) {
    ByteSet s; s.set(b);
    if (f.ci) fold_case(s);
    return bytes_node(s);
  }

  // Parses flags after "(?"; returns true when a scoped group follows (':').
  bool parse_flags(Flags* nf) {
    bool neg = false;
    while (!eof()) {
      int c = peek();
      if (c == ':') { ++i; return true; }
      if (c == ')') { ++i; return false; }
      ++i;
      if (c == '-') { if (neg) fail("repeated flag negation"); neg = true; continue; }
      if (c == 'i') nf->ci = !neg;
      else if (c == 's') nf->dotall = !neg;
      else if (c == 'U') nf->swap_greed = !neg;
      else if (c == 'm') { /* multi-line is always on */ }
      else if (c == 'u') { if (!neg) fail("unicode mode is not supported in byte mode"); }
      else if (c == 'x') fail("verbose mode is not supported");
      else fail("unknown flag");
    }
    fail("unclosed group");
  }

  Ast parse_atom(bool* flag_only) {
    int c = peek();
    switch (c) {
      case '(': {
        ++i;
        Flags saved = f;
        if (peek() == '?') {
          ++i;
          if (peek() == 'P' && peek(1) == '<') i += 2, skip_name();
          else if (peek() == '<') ++i, skip_name();
          else if (peek() == '=' || peek() == '!') fail("look-around is not supported");
          else {
            Flags nf = f;
            if (!parse_flags(&nf)) {  // (?flags) applies to the rest of the group
              f = nf;
              *flag_only = true;
              return Ast{};
            }
            f = nf;
          }
        }
        Ast inner = parse_alt();
        if (peek() != ')') fail("unclosed group");
        ++i;
        f = saved;
        Ast g; g.kind = Ast::GROUP; g.kids.push_back(std::move(inner));
        return g;
      }
      case ')': fail("unopened group");
      case '[': return bytes_node(parse_class());
      case '.': {
        ++i;
        ByteSet s; s.negate();
        if (!f.dotall) s.w['\n' >> 6] &= ~(1ull << ('\n' & 63));
        return bytes_node(s);
      }
      case '^': case '$': {
        ++i;
        Ast a; a.kind = Ast::LOOK; a.look = c == '^' ? LOOK_BOL : LOOK_EOL;  // multi-line: line start / end
        return a;
      }
      case '*': case '+': case '?': case '{': fail("repetition operator missing expression");
      case '\\': {
        ++i;
        const int e = peek();
        if (e == 'b' || e == 'B' || e == 'A' || e == 'z') {
          ++i;
          Ast a; a.kind = Ast::LOOK;
          a.look = e == 'b' ? LOOK_WORDB : e == 'B' ? LOOK_NWORDB : e == 'A' ? LOOK_BOT : LOOK_EOT;
          return a;
        }
        ByteSet s;
        if (parse_escape(&s, false)) return bytes_node(s);
        if (f.ci) fold_case(s);
        return bytes_node(s);
      }
      default:
        ++i;
        return literal_byte(c);
    }
  }

  void skip_name() {
    while (!eof() && peek() != '>') ++i;
    if (eof()) fail("unclosed group name");
    ++i;
  }

  int hexval(int c) {
    if (c >= '0' && c <= '9') return c - '0';
    if (c >= 'a' && c <= 'f') return c - 'a' + 10;
    if (c >= 'A' && c <= 'F') return c - 'A' + 10;
    return -1;
  }

  // Parses the escape after '\'. Fills *s. Returns true when the escape is a
  // class (\d, \w, ...) that must not be case-folded again.
  bool parse_escape(ByteSet* s, bool in_class) {
    if (eof()) fail("incomplete escape");
    int c = peek();
    ++i;
    switch (c) {
      case 'n': s->set('\n'); return false;
      case 't': s->set('\t'); return false;
      case 'r': s->set('\r'); return false;
      case 'f': s->set('\f'); return false;
      case 'v': s->set('\v'); return false;
      case 'a': s->set(7); return false;
      case 'd': *s = digit_set(); return true;
      case 'D': *s = digit_set(); s->negate(); return true;
      case 'w': *s = word_set(); return true;
      case 'W': *s = word_set(); s->negate(); return true;
      case 's': *s = space_set(); return true;
      case 'S': *s = space_set(); s->negate(); return true;
      case 'x': {
        long v = 0;
        if (peek() == '{') {
          ++i;
          int nd = 0;
          while (!eof() && peek() != '}') {
            int h = hexval(peek());
            if (h < 0) fail("invalid hex escape");
            v = v * 16 + h; ++i; ++nd;
            if (v > 0xFF) fail("hex escape above 0xFF in byte mode");
          }
          if (eof() || nd == 0) fail("invalid hex escape");
          ++i;
        } else {
          int h1 = hexval(peek()), h2 = hexval(peek(1));
          if (h1 < 0 || h2 < 0) fail("invalid hex escape");
          v = h1 * 16 + h2; i += 2;
        }
        s->set((int)v);
        return false;
      }
      case 'b': case 'B': case 'A': case 'z':
        fail("invalid escape in class");
      case '<': case '>':
        fail("start / end of word assertions are not supported");
      case 'p': case 'P': fail("unicode classes are not supported in byte mode");
      default:
        if (c >= '0' && c <= '9') fail("backreferences and octal escapes are not supported");
        if (std::string("\\.+*?()|[]{}^$#&-~ ").find((char)c) != std::string::npos) {
          s->set(c);
          return false;
        }
        fail("unrecognized escape");
    }
  }

  ByteSet posix_class(const std::string& name) {
    if (name == "alnum") { ByteSet s = range_set('0', '9'); s.merge(range_set('A', 'Z')); s.merge(range_set('a', 'z')); return s; }
    if (name == "alpha") { ByteSet s = range_set('A', 'Z'); s.merge(range_set('a', 'z')); return s; }
    if (name == "ascii") return range_set(0, 127);
    if (name == "blank") { ByteSet s; s.set(' '); s.set('\t'); return s; }
    if (name == "cntrl") { ByteSet s = range_set(0, 31); s.set(127); return s; }
    if (name == "digit") return digit_set();
    if (name == "graph") return range_set('!', '~');
    if (name == "lower") return range_set('a', 'z');
    if (name == "print") return range_set(' ', '~');
    if (name == "punct") { ByteSet s = range_set('!', '/'); s.merge(range_set(':', '@')); s.merge(range_set('[', '`')); s.merge(range_set('{', '~')); return s; }
    if (name == "space") return space_set();
    if (name == "upper") return range_set('A', 'Z');
    if (name == "word") return word_set();
    if (name == "xdigit") { ByteSet s = digit_set(); s.merge(range_set('A', 'F')); s.merge(range_set('a', 'f')); return s; }
    fail("unknown POSIX class");
  }

  ByteSet parse_class() {
    ++i;  // '['
    bool neg = false;
    if (peek() == '^') { neg = true; ++i; }
    ByteSet set;
    bool first = true;
    while (true) {
      if (eof()) fail("unclosed character class");
      int c = peek();
      if (c == ']' && !first) { ++i; break; }
      first = false;
      if (c == '[' && peek(1) == ':') {
        size_t close = p.find(":]", i + 2);
        if (close == std::string::npos) fail("unclosed POSIX class");
        std::string name = p.substr(i + 2, close - (i + 2));
        bool pneg = !name.empty() && name[0] == '^';
        if (pneg) name = name.substr(1);
        ByteSet ps = posix_class(name);
        if (pneg) ps.negate();
        set.merge(ps);
        i = close + 2;
        continue;
      }
      if (c == '[') { set.merge(parse_class()); continue; }
      if ((c == '&' && peek(1) == '&') || (c == '-' && peek(1) == '-') || (c == '~' && peek(1) == '~'))
        fail("class set operations are not supported");
      ByteSet lo_set;
      int lo = class_atom(&lo_set);
      if (lo < 0) { set.merge(lo_set); continue; }
      if (peek() == '-' && peek(1) != ']' && peek(1) != -1) {
        ++i;
        ByteSet hi_set;
        int hi = class_atom(&hi_set);
        if (hi < 0) fail("invalid class range");
        if (hi < lo) fail("invalid class range");
        set.set_range(lo, hi);
      } else {
        set.set(lo);
      }
    }
    if (f.ci) fold_case(set);
    if (neg) set.negate();
    return set;
  }

  // One class item: returns the byte, or -1 with *s filled for \d-style sets.
  int class_atom(ByteSet* s) {
    int c = peek();
    if (c == '\\') {
      ++i;
      ByteSet e;
      bool is_class = parse_escape(&e, true);
      if (is_class) { *s = e; return -1; }
      return e.single();
    }
    ++i;
    return c;
  }
};

bool has_rep(const Ast& a) {
  switch (a.kind) {
    case Ast::REPEAT: return true;
    case Ast::CONCAT: case Ast::ALT:
      for (auto& k : a.kids) if (has_rep(k)) return true;
      return false;
    case Ast::GROUP: return has_rep(a.kids[0]);
    default: return false;
  }
}

struct NfaBuilder {
  Nfa& n;
  size_t max_states;
  explicit NfaBuilder(Nfa& nfa, size_t mx) : n(nfa), max_states(mx) {}
  int add(Nfa::Type t, int out = -1, int out1 = -1, int set = -1) {
    if (n.st.size() >= max_states) throw ParseError{"regex NFA exceeds the size limit"};
    n.st.push_back({t, out, out1, set});
    return (int)n.st.size() - 1;
  }
  int set_id(const ByteSet& s) {
    for (size_t k = 0; k < n.sets.size(); ++k) if (n.sets[k] == s) return (int)k;
    n.sets.push_back(s);
    return (int)n.sets.size() - 1;
  }
  // Builds `a` so that it continues at `next`; returns the entry state.
  int build(const Ast& a, int next) {
    switch (a.kind) {
      case Ast::EMPTY: return next;
      case Ast::LOOK: return add(Nfa::LOOK, next, -1, a.look);
      case Ast::BYTES: return add(Nfa::BYTES, next, -1, set_id(a.set));
      case Ast::GROUP: return build(a.kids[0], next);
      case Ast::CONCAT: {
        int cur = next;
        for (size_t k = a.kids.size(); k-- > 0;) cur = build(a.kids[k], cur);
        return cur;
      }
      case Ast::ALT: {
        std::vector<int> entries;
        for (auto& k : a.kids) entries.push_back(build(k, next));
        int cur = entries.back();
        for (size_t k = entries.size() - 1; k-- > 0;) cur = add(Nfa::SPLIT, entries[k], cur);
        return cur;
      }
      case Ast::REPEAT: {
        const Ast& kid = a.kids[0];
        int cur;
        if (a.rmax < 0) {
          int loop = add(Nfa::SPLIT);
          int body = build(kid, loop);
          if (a.greedy) { n.st[loop].out = body; n.st[loop].out1 = next; }
          else { n.st[loop].out = next; n.st[loop].out1 = body; }
          cur = loop;
        } else {
          cur = next;
          for (int k = 0; k < a.rmax - a.rmin; ++k) {
            int body = build(kid, cur);
            cur = a.greedy ? add(Nfa::SPLIT, body, next) : add(Nfa::SPLIT, next, body);
          }
        }
        for (int k = 0; k < a.rmin; ++k) cur = build(kid, cur);
        return cur;
      }
    }
    return next;
  }
};

// Whether assertion `look` holds between a position's preceding context
// (0 text start, 1 '\n', 2 word byte, 3 other byte) and its following one
// (0 text end, else as dfa_ctx_of of the next byte).
bool look_holds(int look, uint32_t prev, uint32_t next) {
  switch (look) {
    case LOOK_BOL: return prev == 0 || prev == 1;
    case LOOK_EOL: return next == 0 || next == 1;
    case LOOK_BOT: return prev == 0;
    case LOOK_EOT: return next == 0;
    case LOOK_WORDB: return (prev == 2) != (next == 2);
    default: return (prev == 2) == (next == 2);
  }
}

// Unordered closure (set semantics) used for factor extraction.
void closure_set(const Nfa& n, std::vector<int>& seeds, std::vector<int>* out, bool* match) {
  std::vector<char> seen(n.st.size(), 0);
  std::vector<int> stack(seeds);
  *match = false;
  while (!stack.empty()) {
    int s = stack.back();
    stack.pop_back();
    if (seen[s]) continue;
    seen[s] = 1;
    const auto& st = n.st[s];
    switch (st.t) {
      case Nfa::SPLIT: stack.push_back(st.out1); stack.push_back(st.out); break;
      case Nfa::EPS: case Nfa::LOOK: stack.push_back(st.out); break;
      case Nfa::MATCH: *match = true; break;
      case Nfa::BYTES: out->push_back(s); break;
    }
  }
  std::sort(out->begin(), out->end());
}

struct VecHash {
  size_t operator()(const std::vector<int>& v) const {
    uint64_t h = 1469598103934665603ull;
    for (int x : v) { h ^= (uint64_t)x; h *= 1099511628211ull; }
    return (size_t)h;
  }
};

}  // namespace

Ast parse_regex(const std::string& pattern, bool case_insensitive) {
  Parser ps(pattern);
  ps.f.ci = case_insensitive;
  Ast a = ps.parse_alt();
  if (!ps.eof()) ps.fail("unopened group");
  return a;
}

bool is_pathological(const Ast& a) {
  switch (a.kind) {
    case Ast::REPEAT: {
      bool outer_unbounded = a.rmax < 0 || a.rmax > 1;
      if (outer_unbounded && has_rep(a.kids[0])) return true;
      return is_pathological(a.kids[0]);
    }
    case Ast::CONCAT: case Ast::ALT:
      for (auto& k : a.kids) if (is_pathological(k)) return true;
      return false;
    case Ast::GROUP: return is_pathological(a.kids[0]);
    default: return false;
  }
}

bool extract_literal(const Ast& a, std::string* out) {
  switch (a.kind) {
    case Ast::BYTES: {
      int b = a.set.single();
      if (b < 0) return false;
      out->push_back((char)b);
      return true;
    }
    case Ast::CONCAT:
      for (auto& k : a.kids) if (!extract_literal(k, out)) return false;
      return true;
    case Ast::GROUP: return extract_literal(a.kids[0], out);
    default: return false;
  }
}

void length_bounds(const Ast& a, long* mn, long* mx) {
  switch (a.kind) {
    case Ast::EMPTY: case Ast::LOOK: *mn = 0; *mx = 0; return;
    case Ast::BYTES: *mn = 1; *mx = 1; return;
    case Ast::GROUP: length_bounds(a.kids[0], mn, mx); return;
    case Ast::CONCAT: {
      long smin = 0, smax = 0;
      for (auto& k : a.kids) {
        long a1, b1;
        length_bounds(k, &a1, &b1);
        smin += a1;
        smax = (smax < 0 || b1 < 0) ? -1 : smax + b1;
      }
      *mn = smin; *mx = smax; return;
    }
    case Ast::ALT: {
      long smin = -1, smax = 0;
      for (auto& k : a.kids) {
        long a1, b1;
        length_bounds(k, &a1, &b1);
        smin = smin < 0 ? a1 : std::min(smin, a1);
        smax = (smax < 0 || b1 < 0) ? -1 : std::max(smax, b1);
      }
      *mn = smin < 0 ? 0 : smin; *mx = smax; return;
    }
    case Ast::REPEAT: {
      long a1, b1;
      length_bounds(a.kids[0], &a1, &b1);
      *mn = a1 * a.rmin;
      if (b1 == 0) *mx = 0;
      else *mx = (a.rmax < 0 || b1 < 0) ? -1 : b1 * a.rmax;
      return;
    }
  }
}

Nfa build_nfa(const Ast& a, size_t max_states) {
  Nfa n;
  NfaBuilder b(n, max_states);
  int m = b.add(Nfa::MATCH);
  n.start = b.build(a, m);
  return n;
}

bool has_look(const Nfa& n) {
  for (const auto& st : n.st)
    if (st.t == Nfa::LOOK) return true;
  return false;
}

AnchoredDfa build_anchored_dfa(const Nfa& n, size_t max_states) {
  AnchoredDfa d;
  const bool look = has_look(n);
  bool uses_prev = false;  // an assertion reads the byte before the position
  for (const auto& st : n.st)
    if (st.t == Nfa::LOOK && st.set != LOOK_EOL && st.set != LOOK_EOT) uses_prev = true;
  // Byte classes: refine the 256 bytes by every byte set of the NFA and, when
  // it has assertions, by '\n' and the word class (so a class fixes the
  // context a byte gives its neighbours).
  std::vector<ByteSet> refine(n.sets);
  if (look) {
    ByteSet nl;
    nl.set('\n');
    refine.push_back(nl);
    refine.push_back(word_set());
  }
  std::vector<int> cls(256, 0);
  int ncls = 1;
  for (const auto& s : refine) {
    std::map<std::pair<int, bool>, int> remap;
    int next = 0;
    std::vector<int> nc(256);
    for (int b = 0; b < 256; ++b) {
      auto key = std::make_pair(cls[b], s.has(b));
      auto it = remap.find(key);
      if (it == remap.end()) it = remap.emplace(key, next++).first;
      nc[b] = it->second;
    }
    cls.swap(nc);
    ncls = next;
  }
  d.nclass = (uint32_t)ncls;
  const uint32_t W = d.nclass + 1;  // row width: the classes, then the end of the text
  std::vector<int> rep(ncls, -1);
  for (int b = 0; b < 256; ++b) {
    d.cmap[b] = (uint8_t)cls[b];
    if (rep[cls[b]] < 0) rep[cls[b]] = b;
  }

  // A state is {context, ordered NFA seeds}: the NFA states reached by the
  // bytes so far, before their epsilon closure, which depends on the bytes on
  // both sides of the position. Without assertions that read the byte before,
  // the context is always 0.
  std::unordered_map<std::vector<int>, uint32_t, VecHash> ids;
  std::vector<std::vector<int>> keys(1);  // keys[0]: the dead state
  auto intern = [&](uint32_t ctx, std::vector<int>&& seeds) -> uint32_t {
    if (seeds.empty()) return 0;
    seeds.insert(seeds.begin(), (int)(uses_prev ? ctx : 0u));
    auto it = ids.find(seeds);
    if (it != ids.end()) return it->second;
    if (keys.size() >= max_states) throw ParseError{"regex DFA exceeds the state limit"};
    const uint32_t id = (uint32_t)keys.size();
    ids.emplace(seeds, id);
    keys.push_back(std::move(seeds));
    return id;
  };
  for (uint32_t c = 0; c < 4; ++c) d.start[c] = intern(c, std::vector<int>{n.start});

  // Priority-ordered epsilon closure of `seeds` at a position between
  // contexts prev and next; stops at the first MATCH (leftmost-first: lower-
  // priority threads are dropped once a match is reached) and reports it.
  std::vector<uint32_t> seen(n.st.size(), 0);
  uint32_t epoch = 0;
  std::vector<int> stack;
  auto closure = [&](const std::vector<int>& key, uint32_t prev, uint32_t next, std::vector<int>* out) -> bool {
    ++epoch;
    for (size_t k = 1; k < key.size(); ++k) {
      stack.assign(1, key[k]);
      while (!stack.empty()) {
        const int s = stack.back();
        stack.pop_back();
        if (seen[s] == epoch) continue;
        seen[s] = epoch;
        const auto& st = n.st[s];
        switch (st.t) {
          case Nfa::SPLIT: stack.push_back(st.out1); stack.push_back(st.out); break;
          case Nfa::EPS: stack.push_back(st.out); break;
          case Nfa::LOOK: if (look_holds(st.set, prev, next)) stack.push_back(st.out); break;
          case Nfa::MATCH: return true;
          case Nfa::BYTES: out->push_back(s); break;
        }
      }
    }
    return false;
  };

  std::vector<uint32_t> table(W, DFA_DEAD);  // the dead state's row
  std::vector<int> cl;
  for (size_t s = 1; s < keys.size(); ++s) {
    const std::vector<int> key = keys[s];  // (keys grows below)
    const uint32_t prev = (uint32_t)key[0];
    const size_t row = table.size();
    table.resize(row + W);
    for (uint32_t c = 0; c < d.nclass; ++c) {
      const int b = rep[c];
      const uint32_t nb = dfa_ctx_of(b);
      cl.clear();
      const bool m = closure(key, prev, nb, &cl);
      std::vector<int> seeds;
      for (int x : cl)
        if (n.sets[n.st[x].set].has(b)) seeds.push_back(n.st[x].out);
      const uint32_t t = intern(nb, std::move(seeds));
      table[row + c] = t | (m ? DFA_MATCH : 0u) | (t == 0 ? DFA_DEAD : 0u);
    }
    cl.clear();
    table[row + d.nclass] = DFA_DEAD | (closure(key, prev, 0, &cl) ? DFA_MATCH : 0u);
  }
  d.nstates = (uint32_t)keys.size();
  d.table = std::move(table);
  for (uint32_t c = 0; c < 4; ++c)
    for (uint32_t x = 0; x < W; ++x)
      if (d.table[(size_t)d.start[c] * W + x] & DFA_MATCH) d.can_empty = true;
  return d;
}

std::vector<std::string> extract_factors(const Nfa& n, int max_len, int max_fanout,
                                         int max_count) {
  struct Item { std::string s; std::vector<int> set; bool match; };
  std::vector<std::string> final_;
  std::vector<Item> level(1);
  {
    std::vector<int> seeds{n.start};
    closure_set(n, seeds, &level[0].set, &level[0].match);
  }
  for (int depth = 0; !level.empty(); ++depth) {
    std::vector<const Item*> expandable;
    for (const auto& it : level) {
      if (it.match || depth == max_len) { final_.push_back(it.s); continue; }
      ByteSet bytes;
      for (int s : it.set) bytes.merge(n.sets[n.st[s].set]);
      // The first byte may fan out widely (e.g. [0-9]+ gives ten 1-byte
      // factors); later bytes only extend while the set stays small.
      int limit = depth == 0 ? 128 : max_fanout;
      if (bytes.count() > limit) { final_.push_back(it.s); continue; }
      expandable.push_back(&it);
    }
    std::vector<Item> children;
    for (const Item* it : expandable) {
      for (int b = 0; b < 256; ++b) {
        std::vector<int> seeds;
        for (int s : it->set)
          if (n.sets[n.st[s].set].has(b)) seeds.push_back(n.st[s].out);
        if (seeds.empty()) continue;
        Item c;
        c.s = it->s + (char)b;
        closure_set(n, seeds, &c.set, &c.match);
        if (c.set.empty() && !c.match) continue;
        children.push_back(std::move(c));
      }
    }
    if (final_.size() + children.size() > (size_t)max_count) {
      for (const Item* it : expandable) final_.push_back(it->s);
      break;
    }
    level = std::move(children);
  }
  return final_;
}

}  // namespace ws
