/*
 * CPU oracle for the warpstate scan path — TEST INFRASTRUCTURE ONLY.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, as the checker. See oracle.h for the contract.
 *
 * Regex matching here is a backtracking VM (Thompson-style program with
 * priority-ordered SPLITs, explored depth-first): the first MATCH found in
 * priority order is the leftmost-first match that the reference's regex
 * crate reports (core/src/regex_engine.rs:7-17, core/src/dfa/scan.rs:433-457).
 * It shares no code with the product's DFA construction.
 */
#include <pthread.h>

#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* FNV-1a window hash and the reference prefilter table                      */
/* ------------------------------------------------------------------------ */

uint32_t or_fnv1a_window(const uint8_t* p, uint64_t len, uint32_t window) {
  /* core/src/pattern/compiler.rs:343-350 */
  uint32_t h = 2166136261u;
  uint64_t w = window ? window : 1;
  for (uint64_t i = 0; i < len && i < w; ++i) {
    h ^= p[i];
    h *= 16777619u;
  }
  return h;
}

static uint32_t next_pow2(uint32_t x) {
  uint32_t r = 1;
  while (r < x) r <<= 1;
  return r;
}

void or_prefilter_sizes(const uint32_t* lens, uint32_t n, uint32_t window, uint32_t* n_buckets,
                        uint32_t* n_entries) {
  /* core/src/literal_prefilter.rs:35-64: bucket_count = next_pow2(2*|group|) */
  uint32_t nb = 0;
  for (uint32_t g = 1; g <= window; ++g) {
    uint32_t cnt = 0;
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t pl = lens[i] < window ? lens[i] : window;
      if (pl < 1) pl = 1;
      if (pl == g) ++cnt;
    }
    if (cnt) nb += next_pow2(cnt * 2);
  }
  *n_buckets = nb;
  *n_entries = n;
}

int or_prefilter_build(const uint32_t* lens, const uint32_t* hashes, uint32_t n, uint32_t window,
                       uint32_t* prefix_meta, uint32_t* bucket_ranges, uint32_t* entries) {
  /* core/src/literal_prefilter.rs:23-88 */
  uint32_t nb_total = 0, ne = 0;
  memset(prefix_meta, 0, sizeof(uint32_t) * 4 * window);
  for (uint32_t g = 1; g <= window; ++g) {
    uint32_t cnt = 0;
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t pl = lens[i] < window ? lens[i] : window;
      if (pl < 1) pl = 1;
      if (pl == g) ++cnt;
    }
    if (!cnt) continue;
    uint32_t bc = next_pow2(cnt * 2);
    uint32_t boff = nb_total;
    /* entries are appended bucket by bucket, each bucket in literal order */
    for (uint32_t b = 0; b < bc; ++b) {
      uint32_t start = ne, c = 0;
      for (uint32_t i = 0; i < n; ++i) {
        uint32_t pl = lens[i] < window ? lens[i] : window;
        if (pl < 1) pl = 1;
        if (pl != g || (hashes[i] & (bc - 1)) != b) continue;
        entries[2 * ne] = hashes[i];
        entries[2 * ne + 1] = i;
        ++ne;
        ++c;
      }
      bucket_ranges[2 * (boff + b)] = start;
      bucket_ranges[2 * (boff + b) + 1] = c;
    }
    nb_total += bc;
    prefix_meta[4 * (g - 1) + 0] = boff;
    prefix_meta[4 * (g - 1) + 1] = bc - 1;
    prefix_meta[4 * (g - 1) + 2] = bc;
    prefix_meta[4 * (g - 1) + 3] = 0;
  }
  return 0;
}

uint32_t or_prefilter_probe(const uint32_t* prefix_meta, uint32_t window,
                            const uint32_t* bucket_ranges, uint32_t n_buckets,
                            const uint32_t* entries, uint32_t prefix_len, uint32_t hash,
                            uint32_t* out, uint32_t cap) {
  /* core/src/literal_prefilter.rs:91-119 */
  if (prefix_len < 1 || prefix_len > window) return 0;
  const uint32_t* m = prefix_meta + 4 * (prefix_len - 1);
  if (m[2] == 0) return 0;
  uint32_t b = m[0] + (hash & m[1]);
  if (b >= n_buckets) return 0;
  uint32_t st = bucket_ranges[2 * b], c = bucket_ranges[2 * b + 1], k = 0;
  for (uint32_t e = st; e < st + c; ++e)
    if (entries[2 * e] == hash) {
      if (k < cap) out[k] = entries[2 * e + 1];
      ++k;
    }
  return k;
}

/* ------------------------------------------------------------------------ */
/* Regex: parser -> AST -> backtracking program                             */
/* ------------------------------------------------------------------------ */

typedef struct { uint32_t w[8]; } bset;
static void bs_set(bset* s, int b) { s->w[b >> 5] |= 1u << (b & 31); }
static int bs_has(const bset* s, int b) { return (s->w[b >> 5] >> (b & 31)) & 1; }
static void bs_range(bset* s, int lo, int hi) { for (int b = lo; b <= hi; ++b) bs_set(s, b); }
static void bs_or(bset* d, const bset* s) { for (int i = 0; i < 8; ++i) d->w[i] |= s->w[i]; }
static void bs_not(bset* s) { for (int i = 0; i < 8; ++i) s->w[i] = ~s->w[i]; }
static int bs_count(const bset* s) {
  int c = 0;
  for (int i = 0; i < 8; ++i) c += __builtin_popcount(s->w[i]);
  return c;
}
static void bs_fold(bset* s) {
  for (int b = 'a'; b <= 'z'; ++b)
    if (bs_has(s, b) || bs_has(s, b - 32)) { bs_set(s, b); bs_set(s, b - 32); }
}

enum { N_EMPTY, N_SET, N_CAT, N_ALT, N_REP, N_GROUP, N_LOOK };
/* Zero-width assertions of regex::bytes with multi_line(true) and
 * unicode(false) (core/src/regex_engine.rs:7-17): ^ and $ at line starts and
 * ends, \A and \z at the text's, \b and \B on the ASCII word class. */
enum { L_BOL, L_EOL, L_BOT, L_EOT, L_WORDB, L_NWORDB };
typedef struct {
  int kind;
  bset set;
  int kid, next; /* first child, next sibling */
  int rmin, rmax, greedy;
} node;

typedef struct {
  const char* p;
  size_t n, i;
  node* nodes;
  int nn, cap;
  int ci, dotall, swapg;
  char err[160];
  int failed;
} parser;

static int pfail(parser* ps, const char* m) {
  if (!ps->failed) snprintf(ps->err, sizeof ps->err, "%s at offset %zu", m, ps->i);
  ps->failed = 1;
  return -1;
}

static int new_node(parser* ps, int kind) {
  if (ps->nn == ps->cap) {
    ps->cap = ps->cap ? ps->cap * 2 : 64;
    ps->nodes = (node*)realloc(ps->nodes, sizeof(node) * ps->cap);
  }
  node* x = &ps->nodes[ps->nn];
  memset(x, 0, sizeof *x);
  x->kind = kind;
  x->kid = x->next = -1;
  return ps->nn++;
}

static int peekc(parser* ps, size_t k) { return ps->i + k < ps->n ? (unsigned char)ps->p[ps->i + k] : -1; }

static int parse_alt(parser* ps);

static int hexv(int c) {
  if (c >= '0' && c <= '9') return c - '0';
  if (c >= 'a' && c <= 'f') return c - 'a' + 10;
  if (c >= 'A' && c <= 'F') return c - 'A' + 10;
  return -1;
}

/* Escape after '\'. Returns 1 for class escapes (\d ...), 0 for single bytes,
 * -1 on error. */
static int parse_esc(parser* ps, bset* s) {
  int c = peekc(ps, 0);
  if (c < 0) return pfail(ps, "incomplete escape");
  ps->i++;
  memset(s, 0, sizeof *s);
  switch (c) {
    case 'n': bs_set(s, 10); return 0;
    case 't': bs_set(s, 9); return 0;
    case 'r': bs_set(s, 13); return 0;
    case 'f': bs_set(s, 12); return 0;
    case 'v': bs_set(s, 11); return 0;
    case 'a': bs_set(s, 7); return 0;
    case 'd': bs_range(s, '0', '9'); return 1;
    case 'D': bs_range(s, '0', '9'); bs_not(s); return 1;
    case 'w': case 'W':
      bs_range(s, '0', '9'); bs_range(s, 'A', 'Z'); bs_range(s, 'a', 'z'); bs_set(s, '_');
      if (c == 'W') bs_not(s);
      return 1;
    case 's': case 'S':
      bs_range(s, 9, 13); bs_set(s, ' ');
      if (c == 'S') bs_not(s);
      return 1;
    case 'x': {
      int v = 0;
      if (peekc(ps, 0) == '{') {
        ps->i++;
        int nd = 0;
        while (peekc(ps, 0) != '}' && peekc(ps, 0) >= 0) {
          int h = hexv(peekc(ps, 0));
          if (h < 0) return pfail(ps, "bad hex escape");
          v = v * 16 + h;
          ps->i++;
          if (v > 255 || ++nd > 8) return pfail(ps, "hex escape above 0xFF");
        }
        if (peekc(ps, 0) != '}' || nd == 0) return pfail(ps, "bad hex escape");
        ps->i++;
      } else {
        int a = hexv(peekc(ps, 0)), b = hexv(peekc(ps, 1));
        if (a < 0 || b < 0) return pfail(ps, "bad hex escape");
        v = a * 16 + b;
        ps->i += 2;
      }
      bs_set(s, v);
      return 0;
    }
    default:
      if (strchr("\\.+*?()|[]{}^$#&-~ ", c)) { bs_set(s, c); return 0; }
      return pfail(ps, "unsupported escape");
  }
}

static int posix(const char* nm, size_t len, bset* s) {
  memset(s, 0, sizeof *s);
#define IS(x) (len == strlen(x) && !strncmp(nm, x, len))
  if (IS("alnum")) { bs_range(s, '0', '9'); bs_range(s, 'A', 'Z'); bs_range(s, 'a', 'z'); }
  else if (IS("alpha")) { bs_range(s, 'A', 'Z'); bs_range(s, 'a', 'z'); }
  else if (IS("ascii")) bs_range(s, 0, 127);
  else if (IS("blank")) { bs_set(s, ' '); bs_set(s, 9); }
  else if (IS("cntrl")) { bs_range(s, 0, 31); bs_set(s, 127); }
  else if (IS("digit")) bs_range(s, '0', '9');
  else if (IS("graph")) bs_range(s, 33, 126);
  else if (IS("lower")) bs_range(s, 'a', 'z');
  else if (IS("print")) bs_range(s, 32, 126);
  else if (IS("punct")) { bs_range(s, 33, 47); bs_range(s, 58, 64); bs_range(s, 91, 96); bs_range(s, 123, 126); }
  else if (IS("space")) { bs_range(s, 9, 13); bs_set(s, ' '); }
  else if (IS("upper")) bs_range(s, 'A', 'Z');
  else if (IS("word")) { bs_range(s, '0', '9'); bs_range(s, 'A', 'Z'); bs_range(s, 'a', 'z'); bs_set(s, '_'); }
  else if (IS("xdigit")) { bs_range(s, '0', '9'); bs_range(s, 'A', 'F'); bs_range(s, 'a', 'f'); }
  else return -1;
#undef IS
  return 0;
}

static int parse_class(parser* ps, bset* out) {
  ps->i++; /* '[' */
  int neg = 0, first = 1;
  bset set;
  memset(&set, 0, sizeof set);
  if (peekc(ps, 0) == '^') { neg = 1; ps->i++; }
  for (;;) {
    int c = peekc(ps, 0);
    if (c < 0) return pfail(ps, "unclosed class");
    if (c == ']' && !first) { ps->i++; break; }
    first = 0;
    if (c == '[' && peekc(ps, 1) == ':') {
      const char* st = ps->p + ps->i + 2;
      const char* e = strstr(st, ":]");
      if (!e) return pfail(ps, "unclosed POSIX class");
      int pneg = *st == '^';
      bset ps2;
      if (posix(st + pneg, (size_t)(e - st - pneg), &ps2) < 0) return pfail(ps, "unknown POSIX class");
      if (pneg) bs_not(&ps2);
      bs_or(&set, &ps2);
      ps->i = (size_t)(e - ps->p) + 2;
      continue;
    }
    if (c == '[') {
      bset inner;
      if (parse_class(ps, &inner) < 0) return -1;
      bs_or(&set, &inner);
      continue;
    }
    if ((c == '&' && peekc(ps, 1) == '&') || (c == '-' && peekc(ps, 1) == '-') ||
        (c == '~' && peekc(ps, 1) == '~'))
      return pfail(ps, "class set operations unsupported");
    int lo;
    if (c == '\\') {
      ps->i++;
      bset e;
      int r = parse_esc(ps, &e);
      if (r < 0) return -1;
      if (r == 1) { bs_or(&set, &e); continue; }
      lo = -1;
      for (int b = 0; b < 256; ++b) if (bs_has(&e, b)) lo = b;
    } else {
      lo = c;
      ps->i++;
    }
    if (peekc(ps, 0) == '-' && peekc(ps, 1) != ']' && peekc(ps, 1) >= 0) {
      ps->i++;
      int hi, c2 = peekc(ps, 0);
      if (c2 == '\\') {
        ps->i++;
        bset e;
        int r = parse_esc(ps, &e);
        if (r != 0) return pfail(ps, "bad class range");
        hi = -1;
        for (int b = 0; b < 256; ++b) if (bs_has(&e, b)) hi = b;
      } else {
        hi = c2;
        ps->i++;
      }
      if (hi < lo) return pfail(ps, "bad class range");
      bs_range(&set, lo, hi);
    } else {
      bs_set(&set, lo);
    }
  }
  if (ps->ci) bs_fold(&set);
  if (neg) bs_not(&set);
  *out = set;
  return 0;
}

static int parse_atom(parser* ps, int* flag_only) {
  int c = peekc(ps, 0);
  *flag_only = 0;
  if (c == '(') {
    ps->i++;
    int sci = ps->ci, sdot = ps->dotall, sug = ps->swapg;
    if (peekc(ps, 0) == '?') {
      ps->i++;
      int named = 0;
      if (peekc(ps, 0) == 'P' && peekc(ps, 1) == '<') { ps->i += 2; named = 1; }
      else if (peekc(ps, 0) == '<' && peekc(ps, 1) != '=' && peekc(ps, 1) != '!') { ps->i += 1; named = 1; }
      if (named) {
        while (peekc(ps, 0) >= 0 && peekc(ps, 0) != '>') ps->i++;
        if (peekc(ps, 0) != '>') return pfail(ps, "unclosed group name");
        ps->i++;
      } else if (peekc(ps, 0) == '=' || peekc(ps, 0) == '!') {
        return pfail(ps, "look-around unsupported");
      } else {
        int neg = 0, scoped = 0;
        for (;;) {
          int f = peekc(ps, 0);
          if (f < 0) return pfail(ps, "unclosed group");
          ps->i++;
          if (f == ':') { scoped = 1; break; }
          if (f == ')') break;
          if (f == '-') { neg = 1; continue; }
          if (f == 'i') ps->ci = !neg;
          else if (f == 's') ps->dotall = !neg;
          else if (f == 'U') ps->swapg = !neg;
          else if (f == 'm') {}
          else return pfail(ps, "unsupported flag");
        }
        if (!scoped) { *flag_only = 1; return -2; }
      }
    }
    int inner = parse_alt(ps);
    if (inner < 0 && ps->failed) return -1;
    if (peekc(ps, 0) != ')') return pfail(ps, "unclosed group");
    ps->i++;
    ps->ci = sci; ps->dotall = sdot; ps->swapg = sug;
    int g = new_node(ps, N_GROUP);
    ps->nodes[g].kid = inner;
    return g;
  }
  if (c == ')') return pfail(ps, "unopened group");
  if (c == '^' || c == '$') {
    ps->i++;
    int x = new_node(ps, N_LOOK);
    ps->nodes[x].rmin = c == '^' ? L_BOL : L_EOL;
    return x;
  }
  if (c == '\\') {
    int e = peekc(ps, 1), look = -1;
    if (e == 'b') look = L_WORDB;
    else if (e == 'B') look = L_NWORDB;
    else if (e == 'A') look = L_BOT;
    else if (e == 'z') look = L_EOT;
    if (look >= 0) {
      ps->i += 2;
      int x = new_node(ps, N_LOOK);
      ps->nodes[x].rmin = look;
      return x;
    }
  }
  if (c == '*' || c == '+' || c == '?' || c == '{') return pfail(ps, "repetition missing expression");
  int x = new_node(ps, N_SET);
  bset s;
  memset(&s, 0, sizeof s);
  if (c == '[') {
    if (parse_class(ps, &s) < 0) return -1;
  } else if (c == '.') {
    ps->i++;
    bs_not(&s);
    if (!ps->dotall) s.w[0] &= ~(1u << 10);
  } else if (c == '\\') {
    ps->i++;
    int r = parse_esc(ps, &s);
    if (r < 0) return -1;
    if (r == 0 && ps->ci) bs_fold(&s);
  } else {
    ps->i++;
    bs_set(&s, c);
    if (ps->ci) bs_fold(&s);
  }
  ps->nodes[x].set = s;
  return x;
}

static int parse_num(parser* ps, int* v) {
  size_t st = ps->i;
  long x = 0;
  while (peekc(ps, 0) >= '0' && peekc(ps, 0) <= '9') {
    x = x * 10 + (peekc(ps, 0) - '0');
    if (x > 100000) return 0;
    ps->i++;
  }
  *v = (int)x;
  return ps->i > st;
}

static int parse_concat(parser* ps) {
  int head = -1, tail = -1, count = 0;
  while (peekc(ps, 0) >= 0 && peekc(ps, 0) != '|' && peekc(ps, 0) != ')') {
    int fo = 0;
    int a = parse_atom(ps, &fo);
    if (fo) continue;
    if (a < 0) return -1;
    int c = peekc(ps, 0), mn = -1, mx = -1;
    if (ps->nodes[a].kind == N_LOOK && (c == '*' || c == '+' || c == '?' || c == '{'))
      return pfail(ps, "repetition of an assertion");
    if (c == '*') { mn = 0; mx = -1; ps->i++; }
    else if (c == '+') { mn = 1; mx = -1; ps->i++; }
    else if (c == '?') { mn = 0; mx = 1; ps->i++; }
    else if (c == '{') {
      ps->i++;
      if (!parse_num(ps, &mn)) return pfail(ps, "bad counted repetition");
      mx = mn;
      if (peekc(ps, 0) == ',') {
        ps->i++;
        if (peekc(ps, 0) == '}') mx = -1;
        else if (!parse_num(ps, &mx)) return pfail(ps, "bad counted repetition");
      }
      if (peekc(ps, 0) != '}') return pfail(ps, "unclosed counted repetition");
      ps->i++;
      if (mx != -1 && mx < mn) return pfail(ps, "bad repetition range");
      if (mn > 1000 || mx > 1000) return pfail(ps, "repetition too large");
    }
    if (mn >= 0) {
      int greedy = 1;
      if (peekc(ps, 0) == '?') { greedy = 0; ps->i++; }
      if (ps->swapg) greedy = !greedy;
      int n2 = peekc(ps, 0);
      if (n2 == '*' || n2 == '+' || n2 == '?' || n2 == '{') return pfail(ps, "stacked repetition");
      if (mx == 0) {
        a = new_node(ps, N_EMPTY);
      } else if (!(mn == 1 && mx == 1)) {
        int r = new_node(ps, N_REP);
        ps->nodes[r].kid = a;
        ps->nodes[r].rmin = mn;
        ps->nodes[r].rmax = mx;
        ps->nodes[r].greedy = greedy;
        a = r;
      }
    }
    if (head < 0) head = a; else ps->nodes[tail].next = a;
    tail = a;
    ++count;
  }
  if (count == 0) return new_node(ps, N_EMPTY);
  if (count == 1) return head;
  int c = new_node(ps, N_CAT);
  ps->nodes[c].kid = head;
  return c;
}

static int parse_alt(parser* ps) {
  int first = parse_concat(ps);
  if (first < 0) return -1;
  if (peekc(ps, 0) != '|') return first;
  int alt = new_node(ps, N_ALT);
  ps->nodes[alt].kid = first;
  int tail = first;
  while (peekc(ps, 0) == '|') {
    ps->i++;
    int b = parse_concat(ps);
    if (b < 0) return -1;
    ps->nodes[tail].next = b;
    tail = b;
  }
  return alt;
}

static int has_rep(const node* nd, int x) {
  const node* n = &nd[x];
  if (n->kind == N_REP) return 1;
  if (n->kind == N_CAT || n->kind == N_ALT)
    for (int k = n->kid; k >= 0; k = nd[k].next) if (has_rep(nd, k)) return 1;
  if (n->kind == N_GROUP) return has_rep(nd, n->kid);
  return 0;
}

/* core/src/pattern/compiler.rs:381-405 */
static int pathological(const node* nd, int x) {
  const node* n = &nd[x];
  if (n->kind == N_REP) {
    int outer_unbounded = n->rmax < 0 || n->rmax > 1;
    if (outer_unbounded && has_rep(nd, n->kid)) return 1;
    return pathological(nd, n->kid);
  }
  if (n->kind == N_CAT || n->kind == N_ALT) {
    for (int k = n->kid; k >= 0; k = nd[k].next) if (pathological(nd, k)) return 1;
    return 0;
  }
  if (n->kind == N_GROUP) return pathological(nd, n->kid);
  return 0;
}

/* core/src/pattern/compiler.rs:360-375; writes into buf, returns length or -1 */
static long literal_of(const node* nd, int x, uint8_t* buf, long pos, long cap) {
  const node* n = &nd[x];
  if (n->kind == N_SET) {
    if (bs_count(&n->set) != 1) return -1;
    for (int b = 0; b < 256; ++b)
      if (bs_has(&n->set, b)) { if (pos >= cap) return -1; buf[pos] = (uint8_t)b; return pos + 1; }
    return -1;
  }
  if (n->kind == N_CAT) {
    for (int k = n->kid; k >= 0; k = nd[k].next) {
      pos = literal_of(nd, k, buf, pos, cap);
      if (pos < 0) return -1;
    }
    return pos;
  }
  if (n->kind == N_GROUP) return literal_of(nd, n->kid, buf, pos, cap);
  return -1;
}

/* Program. */
enum { I_SET, I_SPLIT, I_JMP, I_MATCH, I_LOOK };
typedef struct { int op, x, y, set; } instr;
typedef struct {
  instr* code;
  int n, cap;
  bset* sets;
  int nsets, scap;
  bset first;    /* bytes that can start a match */
  int can_empty; /* a match may be empty (start anywhere, the text end included) */
} prog;

static int emit(prog* p, int op, int x, int y, int set) {
  if (p->n == p->cap) {
    p->cap = p->cap ? p->cap * 2 : 64;
    p->code = (instr*)realloc(p->code, sizeof(instr) * p->cap);
  }
  p->code[p->n].op = op;
  p->code[p->n].x = x;
  p->code[p->n].y = y;
  p->code[p->n].set = set;
  return p->n++;
}

static int add_set(prog* p, const bset* s) {
  if (p->nsets == p->scap) {
    p->scap = p->scap ? p->scap * 2 : 16;
    p->sets = (bset*)realloc(p->sets, sizeof(bset) * p->scap);
  }
  p->sets[p->nsets] = *s;
  return p->nsets++;
}

static int gen(prog* p, const node* nd, int x) {
  const node* n = &nd[x];
  if (p->n > 2000000) return -1;
  switch (n->kind) {
    case N_EMPTY: return 0;
    case N_LOOK: emit(p, I_LOOK, n->rmin, 0, 0); return 0;
    case N_SET: emit(p, I_SET, 0, 0, add_set(p, &n->set)); return 0;
    case N_GROUP: return gen(p, nd, n->kid);
    case N_CAT:
      for (int k = n->kid; k >= 0; k = nd[k].next) if (gen(p, nd, k) < 0) return -1;
      return 0;
    case N_ALT: {
      int jumps[4096], nj = 0;
      for (int k = n->kid; k >= 0; k = nd[k].next) {
        if (nd[k].next >= 0) {
          int sp = emit(p, I_SPLIT, 0, 0, 0);
          p->code[sp].x = p->n;
          if (gen(p, nd, k) < 0) return -1;
          if (nj < 4096) jumps[nj++] = emit(p, I_JMP, 0, 0, 0);
          else return -1;
          p->code[sp].y = p->n;
        } else {
          if (gen(p, nd, k) < 0) return -1;
        }
      }
      for (int j = 0; j < nj; ++j) p->code[jumps[j]].x = p->n;
      return 0;
    }
    case N_REP: {
      for (int k = 0; k < n->rmin; ++k) if (gen(p, nd, n->kid) < 0) return -1;
      if (n->rmax < 0) {
        int sp = emit(p, I_SPLIT, 0, 0, 0);
        int body = p->n;
        if (gen(p, nd, n->kid) < 0) return -1;
        emit(p, I_JMP, sp, 0, 0);
        int out = p->n;
        p->code[sp].x = n->greedy ? body : out;
        p->code[sp].y = n->greedy ? out : body;
      } else {
        int splits[1024], ns = 0;
        for (int k = n->rmin; k < n->rmax; ++k) {
          int sp = emit(p, I_SPLIT, 0, 0, 0);
          if (ns < 1024) splits[ns++] = sp; else return -1;
          p->code[sp].x = p->n; /* body */
          if (gen(p, nd, n->kid) < 0) return -1;
        }
        for (int k = 0; k < ns; ++k) {
          int body = p->code[splits[k]].x;
          p->code[splits[k]].x = n->greedy ? body : p->n;
          p->code[splits[k]].y = n->greedy ? p->n : body;
        }
      }
      return 0;
    }
  }
  return -1;
}

static void first_bytes(prog* p) {
  /* bytes that an I_SET reachable by epsilon from pc 0 accepts */
  char* seen = (char*)calloc((size_t)p->n + 1, 1);
  int* st = (int*)malloc(sizeof(int) * ((size_t)p->n + 1) * 2);
  int sp = 0;
  memset(&p->first, 0, sizeof p->first);
  st[sp++] = 0;
  while (sp) {
    int pc = st[--sp];
    if (pc >= p->n || seen[pc]) continue;
    seen[pc] = 1;
    instr* in = &p->code[pc];
    if (in->op == I_SET) bs_or(&p->first, &p->sets[in->set]);
    else if (in->op == I_JMP) st[sp++] = in->x;
    else if (in->op == I_SPLIT) { st[sp++] = in->x; st[sp++] = in->y; }
    else if (in->op == I_LOOK) st[sp++] = pc + 1; /* zero width: taken as satisfiable */
    else { p->can_empty = 1; } /* MATCH reachable without a byte */
  }
  free(seen);
  free(st);
}

static int is_word(int b) {
  return (b >= '0' && b <= '9') || (b >= 'A' && b <= 'Z') || (b >= 'a' && b <= 'z') || b == '_';
}

/* An assertion at position pos of d[0, len): the bytes on either side
 * (none at the text's ends) decide it. */
static int look_holds(int kind, const uint8_t* d, uint64_t len, uint64_t pos) {
  const int prev = pos > 0 ? d[pos - 1] : -1, next = pos < len ? d[pos] : -1;
  switch (kind) {
    case L_BOL: return prev < 0 || prev == '\n';
    case L_EOL: return next < 0 || next == '\n';
    case L_BOT: return prev < 0;
    case L_EOT: return next < 0;
    case L_WORDB: return (prev >= 0 && is_word(prev)) != (next >= 0 && is_word(next));
    default: return (prev >= 0 && is_word(prev)) == (next >= 0 && is_word(next));
  }
}

/* Depth-first backtracking in priority order. */
static int64_t bt_run(const prog* p, const uint8_t* d, uint64_t len, uint64_t start) {
  size_t cap = 256, sp = 0;
  struct th { int pc; uint64_t pos; } *stk = (struct th*)malloc(sizeof(struct th) * cap);
  stk[sp].pc = 0;
  stk[sp].pos = start;
  ++sp;
  int64_t res = -1;
  while (sp) {
    struct th t = stk[--sp];
    for (;;) {
      const instr* in = &p->code[t.pc];
      if (in->op == I_MATCH) { res = (int64_t)t.pos; goto done; }
      if (in->op == I_LOOK) {
        if (!look_holds(in->x, d, len, t.pos)) break;
        t.pc++;
      } else if (in->op == I_SET) {
        if (t.pos >= len || !bs_has(&p->sets[in->set], d[t.pos])) break;
        t.pos++;
        t.pc++;
      } else if (in->op == I_JMP) {
        t.pc = in->x;
      } else {
        if (sp == cap) { cap *= 2; stk = (struct th*)realloc(stk, sizeof(struct th) * cap); }
        stk[sp].pc = in->y;
        stk[sp].pos = t.pos;
        ++sp;
        t.pc = in->x;
      }
    }
  }
done:
  free(stk);
  return res;
}

/* ------------------------------------------------------------------------ */
/* Pattern set                                                              */
/* ------------------------------------------------------------------------ */

struct or_set {
  uint32_t npat;
  int ci;
  /* literals */
  uint32_t nlit;
  uint8_t** lit;
  uint32_t* lit_len;
  uint32_t* lit_pid;
  /* literal index: open addressing on the first `key_len` bytes */
  uint32_t key_len, min_lit, max_lit;
  uint32_t tmask;
  uint64_t* tkey;
  uint32_t* thead; /* head of chain (literal index + 1) */
  uint32_t* tnext; /* per literal: next literal index + 1 with the same key */
  /* regexes */
  uint32_t nrx;
  prog* rx;
  uint32_t* rx_pid;
  uint32_t strategy;
};

static uint8_t fold1(uint8_t b) { return (b >= 'A' && b <= 'Z') ? b + 32 : b; }

static uint64_t key_at(const or_set* s, const uint8_t* p) {
  uint64_t k = 0;
  for (uint32_t t = 0; t < s->key_len; ++t) {
    uint8_t b = p[t];
    if (s->ci) b = fold1(b);
    k |= (uint64_t)b << (8 * t);
  }
  return k;
}

static uint64_t mix(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  return k;
}

static void index_literals(or_set* s) {
  s->key_len = s->min_lit < 8 ? s->min_lit : 8;
  uint32_t sz = 16;
  while (sz < s->nlit * 4) sz <<= 1;
  s->tmask = sz - 1;
  s->tkey = (uint64_t*)calloc(sz, sizeof(uint64_t));
  s->thead = (uint32_t*)calloc(sz, sizeof(uint32_t));
  s->tnext = (uint32_t*)calloc(s->nlit ? s->nlit : 1, sizeof(uint32_t));
  /* insert in reverse so each chain lists literals in ascending index */
  for (uint32_t i = s->nlit; i-- > 0;) {
    uint64_t k = key_at(s, s->lit[i]);
    uint32_t h = (uint32_t)mix(k) & s->tmask;
    while (s->thead[h] && s->tkey[h] != k) h = (h + 1) & s->tmask;
    s->tkey[h] = k;
    s->tnext[i] = s->thead[h];
    s->thead[h] = i + 1;
  }
}

static int lit_eq(const or_set* s, const uint8_t* d, uint32_t i) {
  const uint8_t* l = s->lit[i];
  for (uint32_t t = 0; t < s->lit_len[i]; ++t) {
    uint8_t a = d[t], b = l[t];
    if (s->ci) { a = fold1(a); b = fold1(b); }
    if (a != b) return 0;
  }
  return 1;
}

/* First literal (ascending index) with the key of d; 0 when none (+1 coded). */
static uint32_t chain_at(const or_set* s, const uint8_t* d) {
  uint64_t k = key_at(s, d);
  uint32_t h = (uint32_t)mix(k) & s->tmask;
  while (s->thead[h]) {
    if (s->tkey[h] == k) return s->thead[h];
    h = (h + 1) & s->tmask;
  }
  return 0;
}

or_set* or_compile(const or_pattern* pats, uint32_t n, int case_insensitive, int* status,
                   char* err, size_t err_cap) {
  *status = 0;
  if (n == 0) {
    *status = 1;
    snprintf(err, err_cap, "pattern set is empty");
    return NULL;
  }
  or_set* s = (or_set*)calloc(1, sizeof(or_set));
  s->npat = n;
  s->ci = case_insensitive;
  s->lit = (uint8_t**)calloc(n, sizeof(uint8_t*));
  s->lit_len = (uint32_t*)calloc(n, sizeof(uint32_t));
  s->lit_pid = (uint32_t*)calloc(n, sizeof(uint32_t));
  s->rx = (prog*)calloc(n, sizeof(prog));
  s->rx_pid = (uint32_t*)calloc(n, sizeof(uint32_t));
  for (uint32_t i = 0; i < n; ++i) {
    if (pats[i].len == 0) {
      *status = 2;
      snprintf(err, err_cap, "pattern %u is empty", i);
      or_free(s);
      return NULL;
    }
    if (pats[i].kind == 0) {
      s->lit[s->nlit] = (uint8_t*)malloc(pats[i].len);
      memcpy(s->lit[s->nlit], pats[i].bytes, pats[i].len);
      s->lit_len[s->nlit] = (uint32_t)pats[i].len;
      s->lit_pid[s->nlit++] = i;
      continue;
    }
    parser ps;
    memset(&ps, 0, sizeof ps);
    ps.p = (const char*)pats[i].bytes;
    ps.n = pats[i].len;
    int root = parse_alt(&ps);
    if (!ps.failed && ps.i < ps.n) pfail(&ps, "unopened group");
    if (ps.failed || root < 0) {
      *status = 6;
      snprintf(err, err_cap, "regex %u: %s", i, ps.err);
      free(ps.nodes);
      or_free(s);
      return NULL;
    }
    if (pathological(ps.nodes, root)) {
      *status = 12;
      snprintf(err, err_cap, "regex %u is pathological", i);
      free(ps.nodes);
      or_free(s);
      return NULL;
    }
    uint8_t buf[4096];
    long ll = literal_of(ps.nodes, root, buf, 0, sizeof buf);
    if (ll > 0) {
      s->lit[s->nlit] = (uint8_t*)malloc((size_t)ll);
      memcpy(s->lit[s->nlit], buf, (size_t)ll);
      s->lit_len[s->nlit] = (uint32_t)ll;
      s->lit_pid[s->nlit++] = i;
      free(ps.nodes);
      continue;
    }
    prog* p = &s->rx[s->nrx];
    if (gen(p, ps.nodes, root) < 0) {
      *status = 6;
      snprintf(err, err_cap, "regex %u too large", i);
      free(ps.nodes);
      or_free(s);
      return NULL;
    }
    emit(p, I_MATCH, 0, 0, 0);
    first_bytes(p);
    s->rx_pid[s->nrx++] = i;
    free(ps.nodes);
  }
  s->min_lit = 0xFFFFFFFFu;
  s->max_lit = 0;
  for (uint32_t i = 0; i < s->nlit; ++i) {
    if (s->lit_len[i] < s->min_lit) s->min_lit = s->lit_len[i];
    if (s->lit_len[i] > s->max_lit) s->max_lit = s->lit_len[i];
  }
  if (s->nlit) index_literals(s);
  /* ScanStrategy::select (core/src/specialize.rs:101-179) and
   * should_use_hash_scanner (core/src/cpu/mod.rs:225-230). */
  if (s->nlit == 0) s->strategy = 0;
  else if (s->ci || s->nrx) s->strategy = 1;
  else if (s->nlit >= 2 && s->nlit <= 8 && s->max_lit <= 16) s->strategy = 2;
  else if (s->nlit > 5000) s->strategy = 3;
  else s->strategy = 1;
  return s;
}

void or_free(or_set* s) {
  if (!s) return;
  for (uint32_t i = 0; i < s->nlit; ++i) free(s->lit[i]);
  for (uint32_t r = 0; r < s->npat; ++r) {
    free(s->rx[r].code);
    free(s->rx[r].sets);
  }
  free(s->lit); free(s->lit_len); free(s->lit_pid);
  free(s->rx); free(s->rx_pid);
  free(s->tkey); free(s->thead); free(s->tnext);
  free(s);
}

uint32_t or_literal_count(const or_set* s) { return s->nlit; }
uint32_t or_regex_count(const or_set* s) { return s->nrx; }
uint32_t or_strategy(const or_set* s) { return s->strategy; }

/* Growable match list. */
typedef struct { or_match* m; uint64_t n, cap; } mlist;
static void ml_push(mlist* l, uint32_t pid, uint64_t st, uint64_t en) {
  if (l->n == l->cap) {
    l->cap = l->cap ? l->cap * 2 : 1024;
    l->m = (or_match*)realloc(l->m, sizeof(or_match) * l->cap);
  }
  l->m[l->n].pattern_id = pid;
  l->m[l->n].start = (uint32_t)st;
  l->m[l->n].end = (uint32_t)en;
  l->n++;
}

static int cmp_canon(const void* a, const void* b) {
  const or_match* x = (const or_match*)a;
  const or_match* y = (const or_match*)b;
  if (x->start != y->start) return x->start < y->start ? -1 : 1;
  if (x->pattern_id != y->pattern_id) return x->pattern_id < y->pattern_id ? -1 : 1;
  if (x->end != y->end) return x->end < y->end ? -1 : 1;
  return 0;
}

/* Every occurrence of every literal, duplicates included (AC Standard
 * find_overlapping_iter, core/src/cpu/scan.rs:300-349). */
static void lits_overlapping(const or_set* s, const uint8_t* d, uint64_t len, mlist* out) {
  if (!s->nlit || len < s->min_lit) return;
  for (uint64_t pos = 0; pos + s->min_lit <= len; ++pos) {
    for (uint32_t c = chain_at(s, d + pos); c; c = s->tnext[c - 1]) {
      uint32_t i = c - 1;
      if (pos + s->lit_len[i] <= len && lit_eq(s, d + pos, i))
        ml_push(out, s->lit_pid[i], pos, pos + s->lit_len[i]);
    }
  }
}

/* Aho-Corasick LeftmostLongest find_iter (core/src/cpu/scan.rs:250-293):
 * leftmost start, then longest, ties to the earliest-added literal. */
static void lits_leftmost_longest(const or_set* s, const uint8_t* d, uint64_t len, mlist* out) {
  if (!s->nlit) return;
  uint64_t pos = 0;
  while (pos + s->min_lit <= len) {
    int found = 0;
    for (uint64_t st = pos; st + s->min_lit <= len && !found; ++st) {
      int32_t best = -1;
      for (uint32_t c = chain_at(s, d + st); c; c = s->tnext[c - 1]) {
        uint32_t i = c - 1;
        if (st + s->lit_len[i] > len || !lit_eq(s, d + st, i)) continue;
        if (best < 0 || s->lit_len[i] > s->lit_len[best]) best = (int32_t)i;
      }
      if (best >= 0) {
        ml_push(out, s->lit_pid[best], st, st + s->lit_len[best]);
        pos = st + s->lit_len[best];
        found = 1;
      }
    }
    if (!found) break;
  }
}

static int cmp_multi(const void* a, const void* b) {
  /* core/src/specialize.rs:350-355: start asc, end desc, pattern_id asc */
  const or_match* x = (const or_match*)a;
  const or_match* y = (const or_match*)b;
  if (x->start != y->start) return x->start < y->start ? -1 : 1;
  if (x->end != y->end) return x->end > y->end ? -1 : 1;
  if (x->pattern_id != y->pattern_id) return x->pattern_id < y->pattern_id ? -1 : 1;
  return 0;
}

/* MultiMemchr (core/src/specialize.rs:313-373). */
static void lits_multi_memchr(const or_set* s, const uint8_t* d, uint64_t len, mlist* out) {
  mlist all = {0};
  for (uint32_t i = 0; i < s->nlit; ++i) {
    uint64_t L = s->lit_len[i];
    uint64_t pos = 0;
    while (pos + L <= len) {
      /* memmem find_iter: next occurrence at or after pos, then jump past it */
      uint64_t st = pos;
      while (st + L <= len && memcmp(d + st, s->lit[i], L) != 0) ++st;
      if (st + L > len) break;
      ml_push(&all, s->lit_pid[i], st, st + L);
      pos = st + L;
    }
  }
  if (all.n) qsort(all.m, all.n, sizeof(or_match), cmp_multi);
  uint64_t last_end = 0;
  for (uint64_t k = 0; k < all.n; ++k)
    if (all.m[k].start >= last_end) {
      last_end = all.m[k].end;
      ml_push(out, all.m[k].pattern_id, all.m[k].start, all.m[k].end);
    }
  free(all.m);
}

/* HashScanner::scan (core/src/hash_scan.rs:321-363): length groups in
 * ascending order; at the cursor the first group with hits emits every
 * matching id, and the cursor jumps to its end. */
static void lits_hash_shortest(const or_set* s, const uint8_t* d, uint64_t len, mlist* out) {
  uint32_t ng = 0;
  uint32_t* glen = (uint32_t*)malloc(sizeof(uint32_t) * (s->nlit + 1));
  for (uint32_t i = 0; i < s->nlit; ++i) {
    uint32_t L = s->lit_len[i], k;
    for (k = 0; k < ng && glen[k] != L; ++k) {}
    if (k == ng) glen[ng++] = L;
  }
  for (uint32_t a = 1; a < ng; ++a)
    for (uint32_t b = a; b > 0 && glen[b - 1] > glen[b]; --b) {
      uint32_t t = glen[b]; glen[b] = glen[b - 1]; glen[b - 1] = t;
    }
  uint64_t cur = 0;
  while (cur < len) {
    int hit = 0;
    for (uint32_t g = 0; g < ng && !hit; ++g) {
      uint32_t L = glen[g];
      if (cur + L > len || cur + s->min_lit > len) continue;
      for (uint32_t c = chain_at(s, d + cur); c; c = s->tnext[c - 1]) {
        uint32_t i = c - 1;
        if (s->lit_len[i] != L || !lit_eq(s, d + cur, i)) continue;
        ml_push(out, s->lit_pid[i], cur, cur + L);
        hit = 1;
      }
      if (hit) cur += L;
    }
    if (!hit) cur += 1;
  }
  free(glen);
}

/* Leftmost match start at or after `from`: the first st in [from, len] where
 * the anchored leftmost-first run succeeds (positions whose byte cannot start
 * a match are skipped; the text end only when a match may be empty). Returns
 * the start or -1, the end in *e. */
static int64_t leftmost_from(const prog* p, const uint8_t* d, uint64_t len, uint64_t from, int64_t* e) {
  for (uint64_t st = from; st <= len; ++st) {
    if (!p->can_empty && (st == len || !bs_has(&p->first, d[st]))) continue;
    *e = bt_run(p, d, len, st);
    if (*e >= 0) return (int64_t)st;
  }
  return -1;
}

/* find_iter (core/src/dfa/scan.rs:433-457 calls regex::bytes::Regex::find_iter):
 * search from the cursor; an empty match that ends where the previous match
 * ended is skipped by searching again one byte further (the regex crate's
 * iterator rule for overlapping empty matches); the cursor then moves to the
 * match end. Returns the count, writing up to cap matches (out may be NULL). */
static uint64_t find_iter_into(const or_set* s, uint32_t r, const uint8_t* d, uint64_t len, mlist* ml,
                               or_match* out, uint64_t cap) {
  const prog* p = &s->rx[r];
  uint64_t pos = 0, k = 0, last_end = 0;
  int have_last = 0;
  while (pos <= len) {
    int64_t e = -1;
    int64_t st = leftmost_from(p, d, len, pos, &e);
    if (st < 0) break;
    if (e == st && have_last && (uint64_t)e == last_end) {
      if (pos + 1 > len) break;
      st = leftmost_from(p, d, len, pos + 1, &e);
      if (st < 0) break;
    }
    if (ml) ml_push(ml, s->rx_pid[r], (uint64_t)st, (uint64_t)e);
    if (out && k < cap) {
      out[k].pattern_id = s->rx_pid[r];
      out[k].start = (uint32_t)st;
      out[k].end = (uint32_t)e;
    }
    ++k;
    pos = (uint64_t)e;
    last_end = (uint64_t)e;
    have_last = 1;
  }
  return k;
}

static void regex_find_all(const or_set* s, uint32_t r, const uint8_t* d, uint64_t len, mlist* out) {
  find_iter_into(s, r, d, len, out, NULL, 0);
}

uint64_t or_regex_find_iter(const or_set* s, uint32_t r, const uint8_t* d, uint64_t len,
                            or_match* out, uint64_t cap) {
  return find_iter_into(s, r, d, len, NULL, out, cap);
}

int64_t or_regex_match_at(const or_set* s, uint32_t r, const uint8_t* d, uint64_t len,
                          uint64_t start) {
  return bt_run(&s->rx[r], d, len, start);
}

uint64_t or_scan(const or_set* s, const uint8_t* d, uint64_t len, int mode, or_match* out,
                 uint64_t cap) {
  mlist all = {0};
  if (mode == 0) lits_overlapping(s, d, len, &all);
  else if (s->strategy == 2) lits_multi_memchr(s, d, len, &all);
  else if (s->strategy == 3) lits_hash_shortest(s, d, len, &all);
  else if (s->strategy == 1) lits_leftmost_longest(s, d, len, &all);
  for (uint32_t r = 0; r < s->nrx; ++r) regex_find_all(s, r, d, len, &all);
  if (all.n) qsort(all.m, all.n, sizeof(or_match), cmp_canon);
  uint64_t n = all.n;
  for (uint64_t k = 0; k < n && k < cap; ++k) out[k] = all.m[k];
  free(all.m);
  return n;
}

/* Byte-range split of the overlapping literal scan over `threads` POSIX
 * threads: shard k scans [lo, hi + max_len - 1) and keeps starts in [lo, hi),
 * the halo rule of build_scan_shards (core/src/multi_gpu.rs:170-200). Shards
 * are start-ordered, so their sorted lists concatenate in canonical order.
 * Literal-only sets; returns UINT64_MAX for a set with regexes. */
typedef struct {
  const or_set* s;
  const uint8_t* d;
  uint64_t len, lo, hi;
  mlist out;
} shard_job;

static void* shard_run(void* arg) {
  shard_job* j = (shard_job*)arg;
  uint64_t end = j->hi + (j->s->max_lit ? j->s->max_lit - 1 : 0);
  if (end > j->len) end = j->len;
  mlist tmp = {0};
  lits_overlapping(j->s, j->d + j->lo, end - j->lo, &tmp);
  for (uint64_t k = 0; k < tmp.n; ++k)
    if (tmp.m[k].start < j->hi - j->lo)
      ml_push(&j->out, tmp.m[k].pattern_id, tmp.m[k].start + j->lo, tmp.m[k].end + j->lo);
  free(tmp.m);
  if (j->out.n) qsort(j->out.m, j->out.n, sizeof(or_match), cmp_canon);
  return NULL;
}

uint64_t or_scan_overlapping_mt(const or_set* s, const uint8_t* d, uint64_t len, uint32_t threads,
                                or_match* out, uint64_t cap) {
  if (s->nrx) return UINT64_MAX;
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  shard_job* jobs = (shard_job*)calloc(threads, sizeof(shard_job));
  pthread_t* tid = (pthread_t*)calloc(threads, sizeof(pthread_t));
  const uint64_t step = (len + threads - 1) / threads;
  for (uint32_t t = 0; t < threads; ++t) {
    jobs[t].s = s;
    jobs[t].d = d;
    jobs[t].len = len;
    jobs[t].lo = step * t < len ? step * t : len;
    jobs[t].hi = step * (t + 1) < len ? step * (t + 1) : len;
    pthread_create(&tid[t], NULL, shard_run, &jobs[t]);
  }
  uint64_t n = 0;
  for (uint32_t t = 0; t < threads; ++t) {
    pthread_join(tid[t], NULL);
    for (uint64_t k = 0; k < jobs[t].out.n; ++k, ++n)
      if (n < cap) out[n] = jobs[t].out.m[k];
    free(jobs[t].out.m);
  }
  free(jobs);
  free(tid);
  return n;
}

/* Whole-set scan over `threads` POSIX threads: the literal part as
 * byte-range shards (mode 0, the halo rule above) or one job (mode 1, whose
 * reductions run left to right), plus one job per regex (find_iter is per
 * pattern, core/src/dfa/scan.rs:433-457). Threads take jobs from a shared
 * counter; the lists are concatenated and sorted canonically. Same result as
 * or_scan. */
typedef struct {
  const or_set* s;
  const uint8_t* d;
  uint64_t len;
  int mode;
  uint32_t nshard;
  uint32_t njobs;
  uint32_t next; /* atomic */
  mlist* out;    /* one list per job */
} mt_pool;

static void* pool_run(void* arg) {
  mt_pool* P = (mt_pool*)arg;
  for (;;) {
    uint32_t j = __atomic_fetch_add(&P->next, 1u, __ATOMIC_RELAXED);
    if (j >= P->njobs) return NULL;
    if (j < P->nshard) {
      if (P->mode == 0) {
        const uint64_t step = (P->len + P->nshard - 1) / P->nshard;
        shard_job sj;
        memset(&sj, 0, sizeof sj);
        sj.s = P->s;
        sj.d = P->d;
        sj.len = P->len;
        sj.lo = step * j < P->len ? step * j : P->len;
        sj.hi = step * (j + 1) < P->len ? step * (j + 1) : P->len;
        shard_run(&sj);
        P->out[j] = sj.out;
      } else if (P->s->strategy == 2) {
        lits_multi_memchr(P->s, P->d, P->len, &P->out[j]);
      } else if (P->s->strategy == 3) {
        lits_hash_shortest(P->s, P->d, P->len, &P->out[j]);
      } else if (P->s->strategy == 1) {
        lits_leftmost_longest(P->s, P->d, P->len, &P->out[j]);
      }
    } else {
      regex_find_all(P->s, j - P->nshard, P->d, P->len, &P->out[j]);
    }
  }
}

uint64_t or_scan_mt(const or_set* s, const uint8_t* d, uint64_t len, int mode, uint32_t threads, or_match* out,
                    uint64_t cap) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  mt_pool P;
  memset(&P, 0, sizeof P);
  P.s = s;
  P.d = d;
  P.len = len;
  P.mode = mode;
  P.nshard = s->nlit ? (mode == 0 ? threads : 1) : 0;
  P.njobs = P.nshard + s->nrx;
  P.out = (mlist*)calloc(P.njobs ? P.njobs : 1, sizeof(mlist));
  pthread_t* tid = (pthread_t*)calloc(threads, sizeof(pthread_t));
  for (uint32_t t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, pool_run, &P);
  for (uint32_t t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
  uint64_t total = 0;
  for (uint32_t j = 0; j < P.njobs; ++j) total += P.out[j].n;
  or_match* all = (or_match*)malloc(sizeof(or_match) * (total ? total : 1));
  uint64_t k = 0;
  for (uint32_t j = 0; j < P.njobs; ++j) {
    if (P.out[j].n) memcpy(all + k, P.out[j].m, sizeof(or_match) * P.out[j].n);
    k += P.out[j].n;
    free(P.out[j].m);
  }
  if (total) qsort(all, total, sizeof(or_match), cmp_canon);
  for (uint64_t i = 0; i < total && i < cap; ++i) out[i] = all[i];
  free(all);
  free(P.out);
  free(tid);
  return total;
}
