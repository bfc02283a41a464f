// adp.cpp -- generic ADP grammars: analysis, HIP code generation, hipRTC
// compilation and the device wavefront (declared in adp.h).
//
// Semantics follow the reference's native code generator:
//  * yield analysis, unused-tabulation stripping, evaluation order and rule
//    ids: BaseParsers.scala:95-142 (min/max yields, `rulesOrder`, ids by
//    tabulation name, each tabulation owning `inner.alt` consecutive ids);
//  * alternatives: the normal form of CodeGen.norm (CodeGen.scala:31-61),
//    Or lifted above Concat/Map/Filter, left alternatives first, a
//    concatenation of alternatives numbered left-major (alt = l.alt * r.alt);
//  * cell body: CodeGen.genTab (CodeGen.scala:100-186) -- alternatives in
//    order, split loops ascending with the outer concatenation outermost, a
//    candidate kept only when strictly better (`_c > _c0 || rule == -1`, or
//    against the empty value when one is given), back-pointer = (rule, split
//    positions of the concatenations that need one, left to right);
//  * concatenation index ranges: Concat.apply / CodeGen.gen
//    (BaseParsers.scala:282-345, CodeGen.scala:142-160);
//  * backtrace: CodeGen.genBT / genUnapp (CodeGen.scala:213-252), a BFS
//    queue of (i, j, tabulation id) records rewritten to (i, j, rule, pos);
//  * window: CodeGen.scala:325-348 (best axiom cell (i, i+window) by the
//    signature's h, first one on ties).
//  * sum / count (BaseParsers.scala:145-153,385-392): the cell's value is
//    the sum of its candidates' values (count: their number), a cell with no
//    candidate has none; there is no backtrace (no candidate matches a sum,
//    BaseParsers.scala:150), so those grammars return answers only.
// Not supported (rejected with a message): nested aggregations (an
// aggregate below a concatenation or a map), more than two split positions
// per tabulation.
#include "adp.h"

#include <hip/hiprtc.h>

#include <cstddef>

#include "rna_model.h"

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <vector>

namespace dpx {
namespace adp {

namespace {

// ------------------------------------------------------------------ JSON

struct J {
  enum T { NUL, BOOL, NUM, STR, ARR, OBJ } t = NUL;
  bool b = false;
  double num = 0;
  std::string s;
  std::vector<J> a;
  std::vector<std::pair<std::string, J>> o;
  const J* get(const char* k) const {
    for (auto& kv : o)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct JParser {
  const char* p;
  const char* base;
  std::string err;
  int depth = 0;  // nesting bound: a hostile spec cannot exhaust the stack
  void ws() {
    while (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r') ++p;
  }
  bool fail(const std::string& m) {
    if (err.empty()) err = m + " at offset " + std::to_string(p - base);
    return false;
  }
  bool str(std::string& out) {
    if (*p != '"') return fail("expected string");
    ++p;
    while (*p && *p != '"') {
      if (*p == '\\') {
        ++p;
        switch (*p) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case '/': out += '/'; break;
          case '\\': out += '\\'; break;
          case '"': out += '"'; break;
          case 'u': {
            unsigned v = 0;
            for (int k = 1; k <= 4; ++k) {
              const char c = p[k];
              v = v * 16 + (c >= '0' && c <= '9' ? c - '0' : c >= 'a' && c <= 'f' ? c - 'a' + 10 : c >= 'A' && c <= 'F' ? c - 'A' + 10 : 0);
            }
            if (v >= 128) return fail("non-ASCII \\u escape");
            out += (char)v;
            p += 4;
            break;
          }
          default: return fail("bad escape");
        }
        ++p;
      } else {
        out += *p++;
      }
    }
    if (*p != '"') return fail("unterminated string");
    ++p;
    return true;
  }
  bool val(J& v) {
    ws();
    if (depth > 512) return fail("JSON nested too deeply");
    struct Nest {
      int& d;
      explicit Nest(int& x) : d(x) { ++d; }
      ~Nest() { --d; }
    } nest(depth);
    if (*p == '{') {
      v.t = J::OBJ;
      ++p;
      ws();
      if (*p == '}') { ++p; return true; }
      for (;;) {
        ws();
        std::string k;
        if (!str(k)) return false;
        ws();
        if (*p != ':') return fail("expected ':'");
        ++p;
        J x;
        if (!val(x)) return false;
        v.o.emplace_back(k, std::move(x));
        ws();
        if (*p == ',') { ++p; continue; }
        if (*p == '}') { ++p; return true; }
        return fail("expected ',' or '}'");
      }
    }
    if (*p == '[') {
      v.t = J::ARR;
      ++p;
      ws();
      if (*p == ']') { ++p; return true; }
      for (;;) {
        J x;
        if (!val(x)) return false;
        v.a.push_back(std::move(x));
        ws();
        if (*p == ',') { ++p; continue; }
        if (*p == ']') { ++p; return true; }
        return fail("expected ',' or ']'");
      }
    }
    if (*p == '"') { v.t = J::STR; return str(v.s); }
    if (!strncmp(p, "true", 4)) { v.t = J::BOOL; v.b = true; p += 4; return true; }
    if (!strncmp(p, "false", 5)) { v.t = J::BOOL; p += 5; return true; }
    if (!strncmp(p, "null", 4)) { v.t = J::NUL; p += 4; return true; }
    char* e = nullptr;
    v.num = strtod(p, &e);
    if (e == p) return fail("unexpected character");
    v.t = J::NUM;
    p = e;
    return true;
  }
};

// ------------------------------------------------------------------ model

constexpr int INF = -1;  // unbounded yield (the reference's maxN)
constexpr int MAX_POS = 2;

enum Kind { K_TAB, K_TERM, K_OR, K_MAP, K_FILTER, K_CAT, K_AGG };
enum Term { T_EMPTY, T_EMPTYI, T_EL, T_ELI, T_SEQ, T_EMPTY1, T_EMPTY2, T_EL1, T_EL2 };
enum AggKind { A_MAX, A_MIN, A_MAXBY, A_MINBY, A_SUM, A_COUNT };

struct Agg {
  AggKind k = A_MAX;
  int fn = -1;  // maxBy/minBy key function
  bool operator==(const Agg& o) const { return k == o.k && fn == o.fn; }
};

struct Node {
  Kind k = K_TERM;
  int tab = -1;
  Term term = T_EMPTY;
  int tmin = 0, tmax = 0;
  std::shared_ptr<Node> l, r;
  int fn = -1;
  int track = 0;
  int hint[4] = {0, INF, 0, INF};
  Agg agg;
  // a nested aggregation after flatten: the alternatives of its inner Or
  // (all of them are aggregated together) and which one this copy stands for
  std::vector<std::shared_ptr<Node>> sibs;
  int sib = -1;
};
using NP = std::shared_ptr<Node>;

struct Fn {
  std::string name;
  std::vector<std::string> args;
  std::string body;
};

struct Tab {
  std::string name;
  bool always_valid = false;
  NP p;
  int minv = 0, maxv = 0, id = 0, alt = 1, cat = 0;
  bool used = false;
};

int hm(int a, int b) { return a == INF ? b : b == INF ? a : std::min(a, b); }

struct Model {
  int tracks = 1, window = 0, axiom = -1;
  bool rna = false;  // algebra bodies call the librna energy functions
  std::string alphabet = "char", answer = "int", header, empty;
  int alphabet_bytes = 1, answer_bytes = 4;
  Agg h;
  // sum / count: candidates are added up instead of compared
  bool additive() const { return h.k == A_SUM || h.k == A_COUNT; }
  // integer answers add up to the same value in any order (wrapping), so a
  // cell's candidates may be split over lanes; other types keep the
  // sequential order on one lane
  bool any_order() const {
    static const char* const ints[] = {"int", "unsigned", "unsigned int", "long", "long int", "long long", "unsigned long",
                                       "unsigned long long", "short", "unsigned short", "char", "signed char",
                                       "unsigned char", "int8_t", "int16_t", "int32_t", "int64_t", "uint8_t", "uint16_t",
                                       "uint32_t", "uint64_t", "size_t"};
    if (!additive()) return true;
    for (const char* t : ints)
      if (answer == t) return true;
    return false;
  }
  std::vector<Fn> fns;
  std::map<std::string, int> fn_ix;
  std::vector<Tab> tabs;
  std::map<std::string, int> tab_ix;
  std::vector<int> order;  // evaluation order within a cell (rulesOrder)
  int n_rules = 0, max_pos = 0;
  bool loops = false;
  bool fixed_refs = true;  // every table reference of a cell body is (i + c, j + c')
  bool reg_band = false;   // two tracks, references (i, j - c) / (i - 1, j - c): the register band kernel

  // yields (BaseParsers.scala: Parser.min/max, Concat hints)
  int nmin(const Node* n) const {
    switch (n->k) {
      case K_TAB: return tabs[n->tab].minv;
      case K_TERM: return n->tmin;
      case K_OR: return std::min(nmin(n->l.get()), nmin(n->r.get()));
      case K_CAT: return std::max(nmin(n->l.get()), n->hint[0]) + std::max(nmin(n->r.get()), n->hint[2]);
      default: return nmin(n->l.get());
    }
  }
  int nmax(const Node* n) const {
    switch (n->k) {
      case K_TAB: return tabs[n->tab].maxv;
      case K_TERM: return n->tmax;
      case K_OR: {
        const int a = nmax(n->l.get()), b = nmax(n->r.get());
        return a == INF || b == INF ? INF : std::max(a, b);
      }
      case K_CAT: {
        const int a = hm(nmax(n->l.get()), n->hint[1]), b = hm(nmax(n->r.get()), n->hint[3]);
        return a == INF || b == INF ? INF : a + b;
      }
      default: return nmax(n->l.get());
    }
  }
  // recursive maximal yield with a depth bound on tabulations (analyze's rmax)
  int rmax(const Node* n, int d) const {
    switch (n->k) {
      case K_TAB: return d == 1 ? INF : rmax(tabs[n->tab].p.get(), d - 1);
      case K_TERM: return n->tmax;
      case K_OR: {
        const int a = rmax(n->l.get(), d);
        if (a == INF) return INF;
        const int b = rmax(n->r.get(), d);
        return b == INF ? INF : std::max(a, b);
      }
      case K_CAT: {
        const int a = rmax(n->l.get(), d);
        if (a == INF) return INF;
        const int b = rmax(n->r.get(), d);
        return b == INF ? INF : a + b;
      }
      default: return rmax(n->l.get(), d);
    }
  }
  int alt(const Node* n) const {
    switch (n->k) {
      case K_TAB: case K_TERM: return 1;
      case K_OR: return alt(n->l.get()) + alt(n->r.get());
      case K_CAT: return alt(n->l.get()) * alt(n->r.get());
      default: return alt(n->l.get());
    }
  }
  bool has_bt(const Node* n) const {  // Concat.hasBt: no split position when a side has a fixed yield
    const int a = nmin(n->l.get()), b = nmax(n->l.get()), c = nmin(n->r.get()), d = nmax(n->r.get());
    if (n->track == 0) return !((a == b && a >= 0) || (c == d && c >= 0));
    if (n->track == 1) return !(a == b && a >= 0);
    return !(c == d && c >= 0);
  }
  int cat(const Node* n) const {
    switch (n->k) {
      case K_TAB: case K_TERM: return 0;
      case K_OR: return std::max(cat(n->l.get()), cat(n->r.get()));
      case K_CAT: return cat(n->l.get()) + (has_bt(n) ? 1 : 0) + cat(n->r.get());
      default: return cat(n->l.get());
    }
  }
  void indices(const Node* n, int ix[4]) const {
    ix[0] = std::max(nmin(n->l.get()), n->hint[0]);
    ix[1] = hm(nmax(n->l.get()), n->hint[1]);
    ix[2] = std::max(nmin(n->r.get()), n->hint[2]);
    ix[3] = hm(nmax(n->r.get()), n->hint[3]);
  }
};

bool same(const Node* a, const Node* b) {  // structural equality (the reference compares case classes)
  if (a->k != b->k) return false;
  switch (a->k) {
    case K_TAB: return a->tab == b->tab;
    case K_TERM: return a->term != T_SEQ && a->term == b->term;  // seq() builds a new parser each call
    case K_OR: return same(a->l.get(), b->l.get()) && same(a->r.get(), b->r.get());
    case K_MAP: case K_FILTER: return a->fn == b->fn && same(a->l.get(), b->l.get());
    case K_CAT:
      return a->track == b->track && same(a->l.get(), b->l.get()) && same(a->r.get(), b->r.get()) &&
             !memcmp(a->hint, b->hint, sizeof a->hint);
    case K_AGG: return a->agg == b->agg && same(a->l.get(), b->l.get());
  }
  return false;
}

// ------------------------------------------------------------------ spec -> model

struct Builder {
  Model& m;
  std::string err;
  explicit Builder(Model& mm) : m(mm) {}
  bool fail(const std::string& s) {
    if (err.empty()) err = s;
    return false;
  }
  bool agg_of(const J* j, Agg& a) {
    if (!j) return fail("missing aggregation");
    if (j->t == J::STR) {
      if (j->s == "max") { a.k = A_MAX; return true; }
      if (j->s == "min") { a.k = A_MIN; return true; }
      if (j->s == "sum") { a.k = A_SUM; return true; }
      if (j->s == "count") { a.k = A_COUNT; return true; }
      return fail("unsupported aggregation '" + j->s + "' (max, min, sum, count, maxBy, minBy)");
    }
    if (j->t == J::OBJ) {
      const J* f = j->get("maxBy");
      a.k = A_MAXBY;
      if (!f) { f = j->get("minBy"); a.k = A_MINBY; }
      if (!f || f->t != J::STR || !m.fn_ix.count(f->s)) return fail("maxBy/minBy needs a known function");
      a.fn = m.fn_ix[f->s];
      return true;
    }
    return fail("bad aggregation");
  }
  int fn_of(const J* j) {
    if (!j || j->t != J::STR || !m.fn_ix.count(j->s)) { fail("unknown function" + std::string(j && j->t == J::STR ? " '" + j->s + "'" : "")); return -1; }
    return m.fn_ix[j->s];
  }
  NP node(const J& j, int depth = 0) {
    if (depth > 200) { fail("parser nesting too deep"); return nullptr; }
    if (j.t != J::OBJ) { fail("parser node must be an object"); return nullptr; }
    auto n = std::make_shared<Node>();
    if (const J* t = j.get("tab")) {
      if (t->t != J::STR || !m.tab_ix.count(t->s)) { fail("unknown tabulation"); return nullptr; }
      n->k = K_TAB;
      n->tab = m.tab_ix[t->s];
      return n;
    }
    if (const J* t = j.get("term")) {
      static const std::map<std::string, Term> terms1 = {{"empty", T_EMPTY}, {"emptyi", T_EMPTYI}, {"el", T_EL}, {"eli", T_ELI}, {"seq", T_SEQ}};
      static const std::map<std::string, Term> terms2 = {{"empty", T_EMPTY}, {"empty1", T_EMPTY1}, {"empty2", T_EMPTY2}, {"el1", T_EL1}, {"el2", T_EL2}, {"eli", T_ELI}, {"seq", T_SEQ}};
      const auto& terms = m.tracks == 1 ? terms1 : terms2;
      if (t->t != J::STR || !terms.count(t->s)) { fail("unknown terminal for a " + std::to_string(m.tracks) + "-track grammar"); return nullptr; }
      n->k = K_TERM;
      n->term = terms.at(t->s);
      switch (n->term) {
        case T_EMPTY: case T_EMPTYI: n->tmin = 0; n->tmax = 0; break;
        case T_EMPTY1: case T_EMPTY2: n->tmin = 0; n->tmax = INF; break;
        case T_EL: case T_ELI: case T_EL1: case T_EL2: n->tmin = 1; n->tmax = 1; break;
        case T_SEQ: {
          const J* a = j.get("min");
          const J* b = j.get("max");
          n->tmin = a ? (int)a->num : 1;
          n->tmax = b && b->t == J::NUM ? (int)b->num : INF;
          if (n->tmin < 0 || (n->tmax != INF && n->tmax < n->tmin)) { fail("bad seq bounds"); return nullptr; }
          break;
        }
      }
      return n;
    }
    if (const J* t = j.get("or")) {
      if (t->t != J::ARR || t->a.size() < 2) { fail("'or' needs two or more parsers"); return nullptr; }
      NP acc = node(t->a[0], depth + 1);  // left-nested like Scala's left-associative |
      for (size_t q = 1; q < t->a.size() && acc; ++q) {
        auto o = std::make_shared<Node>();
        o->k = K_OR;
        o->l = acc;
        o->r = node(t->a[q], depth + 1);
        if (!o->r) return nullptr;
        acc = o;
      }
      return acc;
    }
    if (const J* t = j.get("cat")) {
      if (t->t != J::ARR || t->a.size() != 2) { fail("'cat' needs two parsers"); return nullptr; }
      n->k = K_CAT;
      n->l = node(t->a[0], depth + 1);
      n->r = node(t->a[1], depth + 1);
      if (!n->l || !n->r) return nullptr;
      const J* tr = j.get("track");
      n->track = tr ? (int)tr->num : 0;
      if (n->track < 0 || n->track > 2 || (m.tracks == 1) != (n->track == 0)) { fail("concatenation track does not fit the grammar"); return nullptr; }
      if (const J* h = j.get("hint")) {
        if (h->t != J::ARR || h->a.size() != 4) { fail("hint needs [minl,maxl,minr,maxr]"); return nullptr; }
        for (int q = 0; q < 4; ++q) n->hint[q] = h->a[q].t == J::NUM ? (int)h->a[q].num : INF;
      } else if (n->track == 0 && same(n->l.get(), n->r.get())) {
        n->hint[0] = 1;  // Concat.hint: a parser concatenated with itself yields at least 1 on each side
        n->hint[2] = 1;
      }
      return n;
    }
    if (const J* t = j.get("map")) {
      n->k = K_MAP;
      n->l = node(*t, depth + 1);
      n->fn = fn_of(j.get("f"));
      return n->l && n->fn >= 0 ? n : nullptr;
    }
    if (const J* t = j.get("filter")) {
      n->k = K_FILTER;
      n->l = node(*t, depth + 1);
      n->fn = fn_of(j.get("f"));
      return n->l && n->fn >= 0 ? n : nullptr;
    }
    if (const J* t = j.get("agg")) {
      n->k = K_AGG;
      n->l = node(*t, depth + 1);
      if (!n->l || !agg_of(j.get("h"), n->agg)) return nullptr;
      return n;
    }
    fail("unknown parser node");
    return nullptr;
  }
  bool build(const J& root) {
    if (root.t != J::OBJ) return fail("spec must be a JSON object");
    auto num = [&](const char* k, int dflt) { const J* v = root.get(k); return v && v->t == J::NUM ? (int)v->num : dflt; };
    auto strv = [&](const char* k, const std::string& dflt) { const J* v = root.get(k); return v && v->t == J::STR ? v->s : dflt; };
    m.tracks = num("tracks", 1);
    if (m.tracks != 1 && m.tracks != 2) return fail("tracks must be 1 or 2");
    m.window = num("window", 0);
    if (m.window < 0 || (m.window > 0 && m.tracks != 1)) return fail("window applies to single-track grammars");
    if (const J* r = root.get("rna")) m.rna = r->t == J::BOOL && r->b;
    if (m.rna && m.tracks != 1) return fail("librna functions apply to single-track grammars");
    m.alphabet = strv("alphabet", "char");
    m.answer = strv("answer", "int");
    m.alphabet_bytes = num("alphabet_bytes", 1);
    m.answer_bytes = num("answer_bytes", 4);
    if (m.alphabet_bytes <= 0 || m.answer_bytes <= 0 || m.answer_bytes > 256) return fail("bad type sizes");
    m.header = strv("header", "");
    if (const J* e = root.get("empty"); e && e->t == J::STR) m.empty = e->s;
    const J* fns = root.get("functions");
    if (fns && fns->t == J::OBJ) {
      for (auto& kv : fns->o) {
        Fn f;
        f.name = kv.first;
        const J* a = kv.second.get("args");
        const J* b = kv.second.get("body");
        if (!b || b->t != J::STR) return fail("function '" + f.name + "' needs a body");
        if (a && a->t == J::STR) {
          std::stringstream ss(a->s);
          std::string x;
          while (std::getline(ss, x, ',')) {
            x.erase(0, x.find_first_not_of(" \t"));
            x.erase(x.find_last_not_of(" \t") + 1);
            if (!x.empty()) f.args.push_back(x);
          }
        }
        f.body = b->s;
        m.fn_ix[f.name] = (int)m.fns.size();
        m.fns.push_back(f);
      }
    }
    if (!agg_of(root.get("h"), m.h)) return false;
    const J* tabs = root.get("tabulations");
    if (!tabs || tabs->t != J::OBJ || tabs->o.empty()) return fail("no tabulations");
    for (auto& kv : tabs->o) {
      Tab t;
      t.name = kv.first;
      if (const J* v = kv.second.get("always_valid")) t.always_valid = v->t == J::BOOL && v->b;
      if (m.tab_ix.count(t.name)) return fail("duplicate tabulation name");
      m.tab_ix[t.name] = (int)m.tabs.size();
      m.tabs.push_back(t);
    }
    for (auto& kv : tabs->o) {
      const J* p = kv.second.get("p");
      if (!p) return fail("tabulation '" + kv.first + "' has no parser");
      NP n = node(*p);
      if (!n) return false;
      m.tabs[m.tab_ix[kv.first]].p = n;
    }
    const std::string ax = strv("axiom", "");
    if (!m.tab_ix.count(ax)) return fail("axiom must name a tabulation");
    m.axiom = m.tab_ix[ax];
    return true;
  }
};

// ------------------------------------------------------------------ analysis

bool analyze(Model& m, std::string& err) {
  // tabulations reachable from the axiom (analyze strips the others)
  std::function<void(const Node*)> use = [&](const Node* n) {
    if (n->k == K_TAB) {
      Tab& t = m.tabs[n->tab];
      if (t.used) return;
      t.used = true;
      use(t.p.get());
      return;
    }
    if (n->l) use(n->l.get());
    if (n->r) use(n->r.get());
  };
  use(m.tabs[m.axiom].p.get());
  m.tabs[m.axiom].used = true;
  std::vector<int> used;
  for (int t = 0; t < (int)m.tabs.size(); ++t)
    if (m.tabs[t].used) used.push_back(t);
  const int R = (int)used.size();
  // yields: minimum by iteration from above, maximum by bounded unfolding
  for (int t : used) m.tabs[t].minv = 100000;
  for (int it = 0; it < R; ++it)
    for (int t : used) m.tabs[t].minv = m.nmin(m.tabs[t].p.get());
  for (int t : used) m.tabs[t].maxv = m.tabs[t].minv;
  for (int t : used) m.tabs[t].maxv = m.rmax(m.tabs[t].p.get(), R);
  // dependencies within a cell: a concatenation side of yield 0 puts the
  // other side at the same subword
  std::function<void(const Node*, std::set<int>&)> deps = [&](const Node* n, std::set<int>& out) {
    switch (n->k) {
      case K_TAB: out.insert(n->tab); return;
      case K_TERM: return;
      case K_OR: deps(n->l.get(), out); deps(n->r.get(), out); return;
      case K_CAT: {
        int ix[4];
        m.indices(n, ix);
        if (ix[0] == 0) deps(n->r.get(), out);
        if (ix[2] == 0) deps(n->l.get(), out);
        return;
      }
      default: deps(n->l.get(), out); return;
    }
  };
  std::map<int, std::set<int>> dep;
  for (int t : used) deps(m.tabs[t].p.get(), dep[t]);
  std::set<int> done;
  while ((int)done.size() < R) {
    bool any = false;
    for (int t : used) {  // by tabulation index: independent tabulations in spec order
      if (done.count(t)) continue;
      bool ready = true;
      for (int d : dep[t])
        if (!done.count(d)) ready = false;
      if (ready) {
        m.order.push_back(t);
        done.insert(t);
        any = true;
      }
    }
    if (!any) { err = "loop between tabulations within one cell, error in grammar"; return false; }
  }
  // rule ids: tabulations sorted by name, each owning inner.alt ids
  std::vector<int> byname = used;
  std::sort(byname.begin(), byname.end(), [&](int a, int b) { return m.tabs[a].name < m.tabs[b].name; });
  int id = 0;
  for (int t : byname) {
    Tab& T = m.tabs[t];
    T.id = id;
    T.alt = m.alt(T.p.get());
    T.cat = m.cat(T.p.get());
    id += T.alt;
    m.max_pos = std::max(m.max_pos, T.cat);
  }
  m.n_rules = id;
  if (m.max_pos > MAX_POS) { err = "a rule needs more than " + std::to_string(MAX_POS) + " split positions (trace records hold 2)"; return false; }
  if (m.n_rules > 32767) { err = "too many rules"; return false; }
  return true;
}

// alternatives of a tabulation in rule-id order: Or lifted above
// Concat/Map/Filter; aggregations by the signature's h are transparent at
// the top (Aggregate(Aggregate(p,h),h) = Aggregate(p,h)); any other
// aggregation would be nested
bool flatten(const Model& m, const NP& n, bool top, std::vector<NP>& out, std::string& err) {
  switch (n->k) {
    case K_OR: return flatten(m, n->l, top, out, err) && flatten(m, n->r, top, out, err);
    case K_AGG: {
      if (top) {  // the tabulation's own aggregation (CodeGen.scala:31-34)
        if (!(n->agg == m.h)) { err = "a tabulation's aggregation differs from the signature's h"; return false; }
        return flatten(m, n->l, top, out, err);
      }
      // nested: Aggregate(Or(l, r)) = Or(Aggregate(l), Aggregate(r)), as the
      // reference's norm (CodeGen.scala:31-36); each keeps its aggregation
      if (m.additive() || n->agg.k == A_SUM || n->agg.k == A_COUNT) {
        err = "nested sum or count aggregations are not supported";
        return false;
      }
      // The Scala parsers aggregate the union of the inner alternatives
      // (BaseParsers.scala:228, Or.apply = left ++ right) and the winner's
      // rule id carries which alternative won (Concat.apply combines the
      // sides' ids).  So each distributed copy aggregates all of them and
      // yields a candidate only when the winner is its own alternative: one
      // candidate in all, under the rule id the Scala backtrace gives it.
      std::vector<NP> in;
      if (!flatten(m, n->l, false, in, err)) return false;
      for (int q = 0; q < (int)in.size(); ++q) {
        auto c = std::make_shared<Node>(*n);
        c->l = in[q];
        c->sibs = in;
        c->sib = q;
        out.push_back(c);
      }
      return true;
    }
    case K_FILTER: case K_MAP: {
      std::vector<NP> in;
      if (!flatten(m, n->l, top && n->k == K_FILTER, in, err)) return false;
      for (auto& a : in) {
        auto c = std::make_shared<Node>(*n);
        c->l = a;
        out.push_back(c);
      }
      return true;
    }
    case K_CAT: {
      std::vector<NP> L, Rr;
      if (!flatten(m, n->l, false, L, err) || !flatten(m, n->r, false, Rr, err)) return false;
      for (auto& a : L)
        for (auto& b : Rr) {
          auto c = std::make_shared<Node>(*n);
          c->l = a;
          c->r = b;
          out.push_back(c);
        }
      return true;
    }
    default: out.push_back(n); return true;
  }
}

// ------------------------------------------------------------------ code generation

std::string add(const std::string& a, int c) {
  if (c == 0) return a;
  return "(" + a + (c > 0 ? "+" : "-") + std::to_string(c > 0 ? c : -c) + ")";
}

// e == v + c for an index expression built from i, j, integers, + - and
// parentheses (the forms add() and the concatenation rules produce)
bool affine(const std::string& e, char v, int& c) {
  size_t p = 0;
  long ci = 0, cj = 0, k = 0;
  bool ok = true;
  std::function<void(int)> expr = [&](int sign) {
    int sg = sign;
    for (;;) {
      if (p >= e.size()) { ok = false; return; }
      const char ch = e[p];
      if (ch == '(') {
        ++p;
        expr(sg);
        if (p >= e.size() || e[p] != ')') { ok = false; return; }
        ++p;
      } else if (ch == 'i' || ch == 'j') {
        if (p + 1 < e.size() && (isalnum((unsigned char)e[p + 1]) || e[p + 1] == '_')) { ok = false; return; }
        (ch == 'i' ? ci : cj) += sg;
        ++p;
      } else if (isdigit((unsigned char)ch)) {
        long x = 0;
        while (p < e.size() && isdigit((unsigned char)e[p]) && x < (1L << 30)) x = 10 * x + (e[p++] - '0');
        k += sg * x;
      } else {
        ok = false;
        return;
      }
      if (!ok || p >= e.size() || e[p] == ')') return;
      if (e[p] != '+' && e[p] != '-') { ok = false; return; }
      sg = e[p] == '+' ? sign : -sign;
      ++p;
    }
  };
  expr(1);
  if (!ok || p != e.size()) return false;
  c = (int)k;
  return v == 'i' ? ci == 1 && cj == 0 : cj == 1 && ci == 0;
}

// e == v + c for an index expression of names (i, j, split variables k0,
// k1, ...), integers, + - and parentheses, where v is one name
bool lin_is(const std::string& e, const std::string& v, int& c) {
  std::map<std::string, long> co;
  long k = 0;
  size_t p = 0;
  bool ok = true;
  std::function<void(int)> expr = [&](int sign) {
    int sg = sign;
    for (;;) {
      if (p >= e.size()) { ok = false; return; }
      const char ch = e[p];
      if (ch == '(') {
        ++p;
        expr(sg);
        if (!ok || p >= e.size() || e[p] != ')') { ok = false; return; }
        ++p;
      } else if (isalpha((unsigned char)ch) || ch == '_') {
        size_t q = p;
        while (q < e.size() && (isalnum((unsigned char)e[q]) || e[q] == '_')) ++q;
        co[e.substr(p, q - p)] += sg;
        p = q;
      } else if (isdigit((unsigned char)ch)) {
        long x = 0;
        while (p < e.size() && isdigit((unsigned char)e[p]) && x < (1L << 30)) x = 10 * x + (e[p++] - '0');
        k += sg * x;
      } else {
        ok = false;
        return;
      }
      if (!ok || p >= e.size() || e[p] == ')') return;
      if (e[p] != '+' && e[p] != '-') { ok = false; return; }
      sg = e[p] == '+' ? sign : -sign;
      ++p;
    }
  };
  expr(1);
  if (!ok || p != e.size()) return false;
  for (auto& x : co)
    if (x.second != (x.first == v ? 1 : 0)) return false;
  if (!co.count(v)) return false;
  c = (int)k;
  return true;
}

struct Loop {
  std::string var;
  std::vector<std::string> lo, hi;  // var from max(lo) to min(hi), ascending
};

struct AltCode {
  int rule = 0;
  std::vector<Loop> loops;          // outer concatenation first
  std::vector<std::string> arith;   // index conditions
  std::vector<std::string> other;   // filters and VALID, tree order
  std::vector<bool> other_free;     // may be evaluated before the split loops that do not bind its names
  std::vector<std::string> pos;     // split positions, left to right
  std::vector<std::string> var_of_pos;  // loop variable behind each position ("" if an expression)
  struct Kid { int tab; std::string a, b; };
  std::vector<Kid> kids;
  std::string value;
  // nested aggregations (an aggregate below a concatenation or a map, one
  // alternative each after flatten): each is evaluated as a block of its own
  // loops right before the candidate, into _g<id>c (value), _g<id>r (valid)
  // and _g<id>p<q> (its splits, which join the alternative's positions)
  // alts: every alternative of the aggregation's inner Or, in order; sib:
  // the one this alternative stands for (its candidate counts only when the
  // aggregation's winner is that one, _g<id>w == sib)
  struct Nest { std::string id; Agg h; std::vector<std::shared_ptr<AltCode>> alts; int sib = 0; };
  std::vector<Nest> nested;
};

struct Gen {
  const Model& m;
  std::string err;
  int nvar = 0;
  // cell bodies: references to the cell itself read the values the earlier
  // tabulations of the cell left in registers (cv); fixed records whether
  // every reference is (i + c, j + c') -- then no reference reads what the
  // cell writes, and the cell's stores may wait until its last tabulation
  bool cell = false;
  bool fixed = true;
  // with every reference fixed, the reads of other cells become loads at the
  // top of the tabulation (pre: expression -> name, declarations in order),
  // at an index that is in bounds whatever i and j are, so they need no
  // guard and issue together; the rules' conditions still decide their use
  std::map<std::string, std::string>* pre = nullptr;
  std::vector<std::array<std::string, 4>>* pre_decl = nullptr;  // type, name, expression, input "track var c"
  // band kernel with registers: every reference is (i, j - c) or (i - 1,
  // j - c) with c small; it reads R.v<t>_<di>_<dj> / R.r<t>_<di>_<dj>, filled
  // each step from the lanes' histories and the row above the band
  std::set<std::array<int, 3>>* regs = nullptr;  // (tab, -ci, -cj) referenced
  bool reg_ok = true;
  // middle product of a one-track tile: the references of the alternative's
  // split variable v read 32 x 32 blocks staged in LDS -- (i + a, v + b)
  // the left operand L[row][split], (v + a, j + b) the right operand
  // R[split][column] -- and (v + a, v + b) cells of finished tiles; inputs at
  // v + c a staged window of the input.  Any other reference (the cell itself,
  // (i + a, j + b), offsets beyond MIDM) leaves the alternative to the
  // in-tile wavefront (mid_ok false).  need_lo / need_hi: how far inside
  // the row block after the tile's rows and the column block before its
  // columns the split points must stay for every reference to be a finished
  // cell of another tile.
  struct MRef { int t; bool left; int a, b; bool rules; };
  std::vector<MRef>* mrefs = nullptr;
  std::set<int>* minputs = nullptr;  // tracks whose input is staged
  bool mid_ok = true;
  int need_lo = 0, need_hi = 0;
  std::string mvar;  // the split variable (first loop of the alternative)
  static constexpr int MIDM = 16;  // largest offset of a reference from the split variable
  int reg_mref(int t, bool left, int a, int b, bool rule) {
    for (size_t x = 0; x < mrefs->size(); ++x) {
      MRef& r = (*mrefs)[x];
      if (r.t == t && r.left == left && r.a == a && r.b == b) {
        r.rules = r.rules || rule;
        return (int)x;
      }
    }
    mrefs->push_back({t, left, a, b, rule});
    return (int)mrefs->size() - 1;
  }
  std::string mref(int t, const std::string& a, const std::string& b, bool rule) {
    int ca = 0, cb = 0;
    const std::string arr = rule ? "R" : "";
    if (!mvar.empty()) {
      if (lin_is(a, "i", ca) && lin_is(b, mvar, cb) && ca >= 0 && ca <= MIDM && cb >= -MIDM && cb <= MIDM) {
        need_hi = std::max(need_hi, cb);  // column split + b before the tile's columns
        return "sm.mL" + arr + std::to_string(reg_mref(t, true, ca, cb, rule)) + "[_mr * 32 + _mkk]";
      }
      if (lin_is(a, mvar, ca) && lin_is(b, "j", cb) && ca >= -MIDM && ca <= MIDM && cb <= 0 && cb >= -MIDM) {
        need_lo = std::max(need_lo, -ca);  // row split + a below the tile's rows
        return "sm.mR" + arr + std::to_string(reg_mref(t, false, ca, cb, rule)) + "[_mkk * 32 + _mc]";
      }
      if (lin_is(a, mvar, ca) && lin_is(b, mvar, cb) && ca >= -MIDM && cb <= MIDM && ca <= cb) {
        need_hi = std::max(need_hi, cb);  // rows from the tile's rows on, columns before its columns
        return rule ? "((const GLOBAL_ short*)T.r[" + std::to_string(t) + "])[sidx_(z," + a + "," + b + ")]"
                    : "((const GLOBAL_ answer_t*)T.c[" + std::to_string(t) + "])[sidx_(z," + a + "," + b + ")]";
      }
    }
    mid_ok = false;
    return "0";
  }
  explicit Gen(const Model& mm, bool in_cell = false) : m(mm), cell(in_cell) {}
  std::string reg(const char* f, int t, const std::string& a, const std::string& b) {
    int ci = 0, cj = 0;
    if (!affine(a, 'i', ci) || !affine(b, 'j', cj) || ci < -1 || ci > 0 || cj > 0 || -cj + (ci ? 1 : 0) > 4) {
      reg_ok = false;
      return "0";
    }
    regs->insert({t, -ci, -cj});
    return std::string("R.") + f + std::to_string(t) + "_" + std::to_string(-ci) + "_" + std::to_string(-cj);
  }
  std::string hoist(const std::string& type, const std::string& expr, const std::string& in = "") {
    auto it = pre->find(expr);
    if (it != pre->end()) return it->second;
    const std::string nm = "_l" + std::to_string(pre->size());
    (*pre)[expr] = nm;
    pre_decl->push_back({type, nm, expr, in});
    return nm;
  }
  bool self(const std::string& a, const std::string& b) {
    int ca = 0, cb = 0;
    const bool fa = affine(a, 'i', ca), fb = affine(b, 'j', cb);
    if (!fa || !fb) fixed = false;
    return cell && fa && fb && ca == 0 && cb == 0;
  }
  // in-tile form of a one-track cell: other cells through the accessors of
  // the tile's LDS region (which fall back to the table in memory)
  bool tile = false;
  std::set<int>* tile_rules = nullptr;  // tabulations whose rule entries the region needs
  // Reads known to land in the region or a strip index LDS directly (no
  // range tests, no fallback load): set once the region and strips are
  // decided.  t_reg_h: the region's halo (32 for these paths); t_sl / t_sr:
  // tabulations with a left / right strip; t_mid: rules with a middle
  // product and its margins t_ml, t_mh.
  int t_reg_h = -1;
  const std::set<int>* t_sl = nullptr;
  const std::set<int>* t_sr = nullptr;
  const std::set<int>* t_mid = nullptr;
  int t_ml = 0, t_mh = 0;
  AltCode* cur = nullptr;  // the alternative being generated
  // e == base + [lo, hi] (base "i" or "j"), from e itself or the bounds of
  // the alternative's loop variable it is an offset of
  bool span(const std::string& e, const char* base, int& lo, int& hi) const {
    int c = 0;
    if (lin_is(e, base, c)) { lo = hi = c; return true; }
    if (!cur) return false;
    for (auto& L : cur->loops) {
      if (!lin_is(e, L.var, c)) continue;
      bool fl = false, fh = false;
      int l = 0, h = 0, x = 0;
      for (auto& it : L.lo)
        if (lin_is(it, base, x)) { l = fl ? std::max(l, x) : x; fl = true; }
      for (auto& it : L.hi)
        if (lin_is(it, base, x)) { h = fh ? std::min(h, x) : x; fh = true; }
      if (!fl || !fh) return false;
      lo = l + c;
      hi = h + c;
      return true;
    }
    return false;
  }
  std::string tread(int t, const std::string& a, const std::string& b, bool rule) const {
    if (t_reg_h != 32) return "";
    const std::string ts = std::to_string(t), arr = (rule ? "trr_" : "tv_") + ts;
    int alo = 0, ahi = 0, blo = 0, bhi = 0;
    // rows i0 .. i0 + 63, columns j0 - 32 .. j0 + 31 for every value of the loops
    if (span(a, "i", alo, ahi) && span(b, "j", blo, bhi) && alo >= 0 && ahi <= 32 && blo >= -32 && bhi <= 0)
      return arr + "[((" + a + ") - _ri0) * RW_ + ((" + b + ") - _rj0)]";
    if (cur && t_mid && t_mid->count(cur->rule) && cur->loops.size() == 1) {
      // the splits of a rule with a middle product left to the in-tile pass:
      // below KLO (left operand in the left strip, right operand in the
      // region) or above KHI (left in the region, right in the right strip)
      const std::string& v = cur->loops[0].var;
      int ca = 0, cb = 0;
      if (t_sl->count(t) && lin_is(a, "i", ca) && ca == 0 && lin_is(b, v, cb) && cb >= 0 && t_ml + cb <= 32 && t_mh - cb <= 32)
        return arr + "[(" + b + ") >= _rj0 ? (i - _ri0) * RW_ + ((" + b + ") - _rj0) : SLO_" + ts + " + (i - _ri0) * 64 + ((" + b +
               ") - _ri0)]";
      if (t_sr->count(t) && lin_is(a, v, ca) && ca >= 0 && lin_is(b, "j", cb) && cb == 0 && t_ml + ca <= 32 && t_mh - ca <= 32)
        return arr + "[(" + a + ") < _ri0 + RH_ ? ((" + a + ") - _ri0) * RW_ + (j - _rj0) : SRO_" + ts + " + ((" + a +
               ") - _rj0) * 32 + (j - _rj0 - RHALO_)]";
    }
    return "";
  }
  std::string cost(int t, const std::string& a, const std::string& b) {
    if (self(a, b)) return "cv.c" + std::to_string(t);
    if (mrefs) return mref(t, a, b, false);
    if (tile) {
      const std::string f = tread(t, a, b, false);
      int ca = 0, cb = 0;
      // a fixed-offset read of the region is in bounds for any cell of the
      // tile: loaded at the top of the cell with the others
      if (!f.empty() && pre && lin_is(a, "i", ca) && lin_is(b, "j", cb)) return hoist("answer_t", f);
      if (!f.empty()) return f;
      return "rv_" + std::to_string(t) + "(z, T, _ri0, _rj0, " + a + ", " + b + ")";
    }
    if (regs) return reg("v", t, a, b);
    if (pre) return hoist("answer_t", "((const GLOBAL_ answer_t*)T.c[" + std::to_string(t) + "])[sidx_(z," + a + "," + b + ")]");
    return "((const GLOBAL_ answer_t*)T.c[" + std::to_string(t) + "])[idx_(z," + a + "," + b + ")]";
  }
  std::string valid(int t, const std::string& a, const std::string& b) {
    if (!m.empty.empty()) return "!(" + cost(t, a, b) + "==(answer_t)(" + m.empty + "))";
    if (self(a, b)) return "(cv.r" + std::to_string(t) + "!=-1)";
    if (mrefs) return "(" + mref(t, a, b, true) + "!=-1)";
    if (tile) {
      tile_rules->insert(t);
      const std::string f = tread(t, a, b, true);
      int ca = 0, cb = 0;
      if (!f.empty() && pre && lin_is(a, "i", ca) && lin_is(b, "j", cb)) return "(" + hoist("short", f) + "!=-1)";
      if (!f.empty()) return "(" + f + "!=-1)";
      return "(rr_" + std::to_string(t) + "(z, T, _ri0, _rj0, " + a + ", " + b + ")!=-1)";
    }
    if (regs) return "(" + reg("r", t, a, b) + "!=-1)";
    if (pre) return "(" + hoist("short", "((const GLOBAL_ short*)T.r[" + std::to_string(t) + "])[sidx_(z," + a + "," + b + ")]") + "!=-1)";
    return "(((const GLOBAL_ short*)T.r[" + std::to_string(t) + "])[idx_(z," + a + "," + b + ")]!=-1)";
  }
  void scs(int mn, int mx, const std::string& a, const std::string& b, AltCode& o) {
    if (mn == mx) { o.arith.push_back(add(a, mn) + "==" + b); return; }
    if (mx != INF) o.arith.push_back(add(b, -mx) + "<=" + a);
    if (mn > 0) o.arith.push_back(add(a, mn) + "<=" + b);
  }
  // an input element; hoisted (index clamped into the input and its padding)
  // when the index is fixed
  std::string input(int track, const std::string& a) {
    int c = 0;
    const std::string in = track == 1 ? "_in1" : "_in2";
    char v = 'i';
    if (minputs && !mvar.empty() && lin_is(a, mvar, c) && c >= -MIDM && c <= MIDM) {
      minputs->insert(track);
      return "sm.mI" + std::to_string(track) + "[_mkk + " + std::to_string(c + MIDM) + "]";
    }
    if (pre && (affine(a, 'i', c) || (v = 'j', affine(a, 'j', c))))
      return hoist("input_t", in + "[min(max(" + a + ", 0), z.n" + std::to_string(m.tracks == 1 ? 1 : track) + ")]",
                   std::to_string(track) + " " + v + " " + std::to_string(c));
    return in + "[" + a + "]";
  }
  std::string call(int fn, const std::vector<std::string>& args) const {
    std::string s = "f_" + m.fns[fn].name + "(_x";
    for (auto& x : args) s += "," + x;
    return s + ")";
  }
  int nagg = 0;  // nested aggregations generated for this alternative
  bool gen(const Node* n, const std::string& a, const std::string& b, AltCode& o, std::vector<std::string>& vals) {
    cur = &o;
    switch (n->k) {
      case K_AGG: {
        // a nested aggregation (CodeGen.scala:108-135): its one alternative
        // (flatten distributed it over Or) aggregated into a local best; the
        // candidate needs one, and the best one's splits join the
        // alternative's positions (genUnapp passes through, CodeGen.scala:218)
        if (mrefs) mid_ok = false;
        AltCode::Nest nest;
        nest.id = std::to_string(nagg++);
        nest.h = n->agg;
        nest.sib = n->sib < 0 ? 0 : n->sib;
        const std::vector<NP> one{n->l};
        for (auto& sn : n->sib < 0 ? one : n->sibs) {
          auto alt = std::make_shared<AltCode>();
          alt->rule = o.rule;
          std::vector<std::string> iv;
          if (!gen(sn.get(), a, b, *alt, iv)) return false;
          cur = &o;
          if (iv.size() != 1) {
            err = "a nested aggregation's alternative does not yield one answer (map it with a function)";
            return false;
          }
          alt->value = iv[0];
          nest.alts.push_back(alt);
        }
        const AltCode& in = *nest.alts[nest.sib];
        for (size_t q = 0; q < in.pos.size(); ++q) {
          o.pos.push_back("_g" + nest.id + "p" + std::to_string(q));
          o.var_of_pos.push_back(in.var_of_pos[q]);
        }
        for (auto& k : in.kids) o.kids.push_back(k);
        vals.push_back("_g" + nest.id + "c");
        o.nested.push_back(nest);
        return true;
      }
      case K_TAB: {
        const Tab& T = m.tabs[n->tab];
        if (m.tracks == 1) scs(T.minv, T.maxv, a, b, o);
        if (!T.always_valid) {
          // guarded, so that it is safe wherever it is evaluated
          const std::string g = m.tracks == 1 ? "(" + a + ")<=(" + b + ") && (" + b + ")<=j && "
                                              : "(" + a + ")>=0 && (" + b + ")>=0 && ";
          o.other.push_back("(" + g + valid(n->tab, a, b) + ")");
          o.other_free.push_back(true);
        }
        o.kids.push_back({n->tab, a, b});
        vals.push_back(cost(n->tab, a, b));
        return true;
      }
      case K_TERM: {
        switch (n->term) {
          case T_EMPTY:
            if (m.tracks == 2) { o.arith.push_back("0==" + a); o.arith.push_back("0==" + b); }
            scs(0, 0, a, b, o);
            return true;  // Unit: no value
          case T_EMPTYI: scs(0, 0, a, b, o); vals.push_back("(" + a + ")"); return true;
          case T_EMPTY1: o.arith.push_back("0==" + a); vals.push_back("(" + b + ")"); return true;
          case T_EMPTY2: o.arith.push_back("0==" + b); vals.push_back("(" + a + ")"); return true;
          case T_EL: case T_EL1: scs(1, 1, a, b, o); vals.push_back(input(1, a)); return true;
          case T_EL2: scs(1, 1, a, b, o); vals.push_back(input(2, a)); return true;
          case T_ELI: scs(1, 1, a, b, o); vals.push_back("(" + a + ")"); return true;
          case T_SEQ: scs(n->tmin, n->tmax, a, b, o); vals.push_back("T2ii{" + a + "," + b + "}"); return true;
        }
        return true;
      }
      case K_FILTER:
        o.other.push_back(call(n->fn, {a, b}));
        o.other_free.push_back(a == "i" && b == "j");  // a filter of the cell itself
        return gen(n->l.get(), a, b, o, vals);
      case K_MAP: {
        std::vector<std::string> in;
        if (!gen(n->l.get(), a, b, o, in)) return false;
        vals.push_back(call(n->fn, in));
        return true;
      }
      case K_CAT: {
        int ix[4];
        m.indices(n, ix);
        const bool bt = m.has_bt(n);
        std::string k, var;
        std::string la = a, lb = b, ra = a, rb = b;  // subwords of the two sides
        if (n->track == 0) {
          const int iL = ix[0], iU = ix[1], jL = ix[2], jU = ix[3];
          if (iL == iU && jL == jU) {
            o.arith.push_back(add(a, iL + jL) + "==" + b);
            k = add(a, iL);
          } else if (iL == iU) {
            o.arith.push_back(add(a, iL + jL) + "<=" + b);
            if (jU != INF) o.arith.push_back(add(b, -iL - jU) + "<=" + a);
            k = add(a, iL);
          } else if (jL == jU) {
            o.arith.push_back(add(a, jL + iL) + "<=" + b);
            if (iU != INF) o.arith.push_back(add(b, -jL - iU) + "<=" + a);
            k = add(b, -jL);
          } else if (!bt) {
            // a side of fixed yield whose hint excludes it (lower bound above
            // upper): the reference's loop has an empty range
            o.arith.push_back("0==1");
            k = add(a, iL);
          } else {
            var = "k" + std::to_string(nvar++);
            if (mvar.empty()) mvar = var;
            Loop L;
            L.var = var;
            L.lo.push_back(add(a, iL));
            if (jU != INF) L.lo.push_back(add(b, -jU));
            L.hi.push_back(add(b, -jL));
            if (iU != INF) L.hi.push_back(add(a, iU));
            o.loops.push_back(L);
            k = var;
          }
          lb = k;
          ra = k;
        } else {
          // two tracks: -~ splits track 1 (left parses in1[k..i), right the
          // cell (k, j)); ~- splits track 2 (left the cell (i, k), right in2[k..j))
          const int l = n->track == 1 ? ix[0] : ix[2], u = n->track == 1 ? ix[1] : ix[3];
          const std::string& v = n->track == 1 ? a : b;
          if (l == u) {
            o.arith.push_back(std::to_string(l) + "<=" + v);
            k = add(v, -l);
          } else {
            var = "k" + std::to_string(nvar++);
            Loop L;
            L.var = var;
            L.lo.push_back("0");
            if (u != INF) L.lo.push_back(add(v, -u));  // left yield <= u (BaseParsers.scala:296-299)
            L.hi.push_back(add(v, -l));
            o.loops.push_back(L);
            k = var;
          }
          if (n->track == 1) { la = k; lb = a; ra = k; rb = b; }
          else { la = a; lb = k; ra = k; rb = b; }
        }
        if (!var.empty() && !bt) { err = "internal: split loop without a back-pointer"; return false; }
        if (!gen(n->l.get(), la, lb, o, vals)) return false;
        if (bt) { o.pos.push_back(k); o.var_of_pos.push_back(var); }
        return gen(n->r.get(), ra, rb, o, vals);
      }
      default:
        err = "internal: unexpected node";
        return false;
    }
  }
};

std::string join(const std::vector<std::string>& v, const char* sep) {
  std::string s;
  for (size_t q = 0; q < v.size(); ++q) s += (q ? sep : "") + v[q];
  return s;
}

// a left term of the middle product (its generated expression and the split
// variable it names), computed once per staged block
struct MidHoist { std::string expr, var; };

// the top-level arguments of a generated call "f_name(a, b, ...)" ((), [] and
// {} nest); false when e is not one call
bool split_call(const std::string& e, std::string& head, std::vector<std::string>& args) {
  if (e.compare(0, 2, "f_") != 0 || e.back() != ')') return false;
  const size_t lp = e.find('(');
  if (lp == std::string::npos) return false;
  head = e.substr(0, lp);
  for (char c : head)
    if (!isalnum((unsigned char)c) && c != '_') return false;
  args.clear();
  int depth = 0;
  size_t st = lp + 1;
  for (size_t q = lp + 1; q < e.size(); ++q) {
    const char c = e[q];
    if (c == '(' || c == '[' || c == '{') {
      ++depth;
    } else if (c == ')' || c == ']' || c == '}') {
      if (depth == 0) {
        if (q != e.size() - 1) return false;
        args.push_back(e.substr(st, q - st));
        return true;
      }
      --depth;
    } else if (c == ',' && depth == 0) {
      args.push_back(e.substr(st, q - st));
      st = q + 1;
    }
  }
  return false;
}

// e names only the row i, the split variable, staged left operands and
// inputs, the context, the algebra functions and constants: no column j, no
// right operand, no other loop variable
bool left_only(const std::string& e, const std::string& var) {
  static const std::set<std::string> ok{"_x", "i", "_mr", "_mkk", "sm", "z", "T", "_in1", "_in2", "T2ii", "answer_t", "input_t",
                                        "min", "max", "sidx_", "GLOBAL_", "const", "short", "int", "long", "char", "unsigned"};
  if (e.find("sm.mR") != std::string::npos) return false;
  for (size_t q = 0; q < e.size();) {
    const char c = e[q];
    if (isalpha((unsigned char)c) || c == '_') {
      size_t r = q;
      while (r < e.size() && (isalnum((unsigned char)e[r]) || e[r] == '_')) ++r;
      const std::string id = e.substr(q, r - q);
      const bool member = q > 0 && e[q - 1] == '.';
      if (!member && !ok.count(id) && id != var && id.compare(0, 2, "f_") != 0) return false;
      q = r;
    } else if (isdigit((unsigned char)c)) {
      while (q < e.size() && (isalnum((unsigned char)e[q]) || e[q] == '.')) ++q;  // a number and its suffix
    } else {
      ++q;
    }
  }
  return true;
}

// e with each innermost call that takes a staged left operand directly and is
// a left term replaced by its hoisted value (registered in H)
std::string hoist_left(const std::string& e, const std::string& var, std::vector<MidHoist>& H) {
  std::string head;
  std::vector<std::string> args;
  if (!split_call(e, head, args)) return e;
  bool direct = false;
  for (auto& a : args)
    if (a.compare(0, 5, "sm.mL") == 0 && a.size() > 5 && isdigit((unsigned char)a[5]) && a.back() == ']' && a.find(',') == std::string::npos)
      direct = true;
  if (direct && left_only(e, var)) {
    size_t h = 0;
    while (h < H.size() && !(H[h].expr == e && H[h].var == var)) ++h;
    if (h == H.size()) H.push_back({e, var});
    const std::string hs = std::to_string(h);
    return "mu" + hs + "_<MUH" + hs + "_>(_x, z, T, _in1, _in2, i, " + var + ", _mr, _mkk, sm)";
  }
  std::string r = head + "(";
  for (size_t q = 0; q < args.size(); ++q) r += (q ? "," : "") + hoist_left(args[q], var, H);
  return r + ")";
}

std::string bound(const std::vector<std::string>& v, const char* f) {
  std::string s = v[0];
  for (size_t q = 1; q < v.size(); ++q) s = std::string(f) + "(" + s + "," + v[q] + ")";
  return s;
}

// The energy parameters as the device holds them (DevRna, rna_model.h), under
// another name; the offsets are checked against the host's layout.
std::string rna_prelude(const Model& m) {
  std::ostringstream s;
  s << "struct RnaPar {\n"
       "  int stack[8][8];\n"
       "  int mmH[8][5][5], mmI[8][5][5], mm1nI[8][5][5], mm23I[8][5][5], mmM[8][5][5], mmExt[8][5][5];\n"
       "  int d5[8][5], d3[8][5];\n"
       "  int hairpin[31], bulge[31];\n"
       "  int il_ent[61];\n"
       "  short F[31][31];\n"
       "  int ninio0, ninio_max, il5, ml_closing, ml_intern, term_au;\n"
       "  int n_loop[3];\n"
       "  int loop_key[3][40];\n"
       "  int loop_e[3][40];\n"
       "  short int11[8][8][5][5];\n"
       "  short int21[8][8][5][5][5];\n"
       "  short int22[8][8][5][5][5][5];\n"
       "};\n"
       "// the small tables (all but int21 and int22), as RnaPar starts: a copy\n"
       "// in LDS serves the energy functions of the tile kernels\n"
       "struct RnaHead {\n"
       "  int stack[8][8];\n"
       "  int mmH[8][5][5], mmI[8][5][5], mm1nI[8][5][5], mm23I[8][5][5], mmM[8][5][5], mmExt[8][5][5];\n"
       "  int d5[8][5], d3[8][5];\n"
       "  int hairpin[31], bulge[31];\n"
       "  int il_ent[61];\n"
       "  short F[31][31];\n"
       "  int ninio0, ninio_max, il5, ml_closing, ml_intern, term_au;\n"
       "  int n_loop[3];\n"
       "  int loop_key[3][40];\n"
       "  int loop_e[3][40];\n"
       "  short int11[8][8][5][5];\n"
       "};\n"
       "static_assert(__builtin_offsetof(RnaHead, int11) == __builtin_offsetof(RnaPar, int11) && sizeof(RnaHead) % 4 == 0 &&\n"
       "              sizeof(RnaHead) <= __builtin_offsetof(RnaPar, int21), \"energy table head\");\n";
  if (m.rna) {
#define OFF(f) << " && __builtin_offsetof(RnaPar, " #f ") == " << offsetof(DevRna, f)
    s << "static_assert(sizeof(RnaPar) == " << sizeof(DevRna) OFF(mmH) OFF(mmExt) OFF(d5) OFF(hairpin) OFF(bulge) OFF(il_ent)
      OFF(F) OFF(ninio0) OFF(term_au) OFF(n_loop) OFF(loop_key) OFF(loop_e) OFF(int11) OFF(int21) OFF(int22)
      << ", \"energy parameter layout\");\n";
#undef OFF
  }
  return s.str();
}

// The librna functions the reference's RNA algebras call
// (librna_impl.h:106-291), over the context's resident tables: the sequence
// is in1 as codes (A=1, C=2, G=3, U=4, as RNASignature.convert), hl is the
// hairpin entropy by loop size.  Interior loops and bulges up to 30 per side
// (the reference's grammars bound them with seq(1,30)).
std::string rna_functions() {
  return R"(#define NO_BP 7
// pair type without branches: three bits per (x - 1) * 5 + y (CG 1, GC 2,
// GU 3, UG 4, AU 5, UA 6, anything else 7)
__device__ __forceinline__ int bp_index(int x, int y) {
  const unsigned f = (unsigned)((x - 1) * 5 + y);
  return (unsigned)(x - 1) < 4u && (unsigned)y < 5u ? (int)((0xf3eefaff9ffdfffull >> (3 * f)) & 7) : 7;
}
__device__ __forceinline__ int rS_(const Ctx& x, int k) { return IN1V_(x)[k]; }
__device__ __forceinline__ int rBP_(const Ctx& x, int a, int b) { return bp_index(rS_(x, a), rS_(x, b)); }
__device__ __forceinline__ int rTermau_(const Ctx& x, int i, int j) {
  const int p = rS_(x, i), q = rS_(x, j);
  return ((p == 3 && q == 2) || (p == 2 && q == 3)) ? 0 : x.rh->term_au;
}
__device__ __forceinline__ int rSr_(const Ctx& x, int i, int j) { return x.rh->stack[rBP_(x, i, j)][rBP_(x, j - 1, i + 1)]; }
__device__ __forceinline__ int rHl_(const Ctx& x, int i, int j) {
  const int size = j - i - 1;
  if (size < 3) return 600;
  const int cls = size == 3 ? 0 : size == 4 ? 1 : size == 6 ? 2 : -1;
  if (cls >= 0) {
    int key = 0;
    for (int m = 0; m <= size + 1; ++m) key = key * 8 + rS_(x, i + m);
    for (int e = 0; e < x.rh->n_loop[cls]; ++e)
      if (x.rh->loop_key[cls][e] == key) return x.rh->loop_e[cls][e];
  }
  const int ent = x.hl[size];
  if (size == 3) return ent + rTermau_(x, i, j);
  return ent + x.rh->mmH[rBP_(x, i, j)][rS_(x, i + 1)][rS_(x, j - 1)];
}
__device__ __forceinline__ int rBl_(const Ctx& x, int i, int k, int l, int j) {
  const int size = l - k + 1;
  if (size < 0 || size > 30) return 10000000;
  if (size <= 1) return (size ? x.rh->bulge[1] : 0) + x.rh->stack[rBP_(x, i, j)][rBP_(x, j - 1, l + 1)];
  return x.rh->bulge[size] + rTermau_(x, i, j) + rTermau_(x, j - 1, l + 1);
}
__device__ __forceinline__ int rBr_(const Ctx& x, int i, int k, int l, int j) {
  const int size = l - k + 1;
  if (size < 0 || size > 30) return 10000000;
  if (size <= 1) return (size ? x.rh->bulge[1] : 0) + x.rh->stack[rBP_(x, i, j)][rBP_(x, k - 1, i + 1)];
  return x.rh->bulge[size] + rTermau_(x, i, j) + rTermau_(x, k - 1, i + 1);
}
__device__ __forceinline__ int rIl_(const Ctx& x, int i, int k, int l, int j) {
  const LDS_ RnaHead* P = x.rh;  // the small tables, in LDS
  const RnaPar* Q = x.rna;        // int21, int22 in memory
  const int sl = k - i - 1, sr = j - l - 1;
  if (sl <= 0 || sr <= 0 || sl > 30 || sr > 30) {  // bulges, and sizes outside the tables
    if (sl == 0) return rBr_(x, i, l + 1, j - 1, j);
    if (sr == 0) return rBl_(x, i, i + 1, k - 1, j);
    return 10000000;
  }
  const int ocb = rBP_(x, i, j), icb = rBP_(x, l, k);
  const int olb = rS_(x, i + 1), orb = rS_(x, j - 1), ilb = rS_(x, l + 1), irb = rS_(x, k - 1);
  if (sl <= 2 && sr <= 2) {  // 1x1, 1x2, 2x1, 2x2: the loop tables
    if (sl == 1 && sr == 1) return P->int11[ocb][rBP_(x, j - 2, i + 2)][olb][orb];
    if (sl == 1) return Q->int21[ocb][rBP_(x, j - 3, i + 2)][olb][rS_(x, j - 2)][orb];
    if (sr == 1) return Q->int21[rBP_(x, j - 2, i + 3)][ocb][orb][olb][rS_(x, i + 2)];
    return Q->int22[ocb][rBP_(x, j - 3, i + 3)][olb][rS_(x, i + 2)][rS_(x, j - 2)][orb];
  }
  // every other size takes one path (the lanes of a wave differ in sr): the
  // size term, and the outer and inner mismatch from one of mmI (generic),
  // mm1nI (1xn) and mm23I (2x3, with il5 + ninio0 for the size), which lie
  // one after another
  const int t = sl == 1 || sr == 1 ? 1 : (sl == 2 && sr == 3) || (sl == 3 && sr == 2) ? 2 : 0;
  const LDS_ int* mm = (const LDS_ int*)((const LDS_ char*)P + __builtin_offsetof(RnaHead, mmI)) + t * 200;
  const int size = t == 2 ? P->il5 + P->ninio0 : P->F[sl][sr];
  return size + mm[ocb * 25 + olb * 5 + orb] + mm[icb * 25 + ilb * 5 + irb];
}
__device__ __forceinline__ int rExt_(const Ctx& x, int i, int j) {
  const int b = rBP_(x, i, j);
  if (i > 0 && j + 1 < x.n1) return x.rh->mmExt[b][rS_(x, i - 1)][rS_(x, j + 1)];
  if (i > 0) return x.rh->d5[b][rS_(x, i - 1)];
  if (j + 1 < x.n1) return x.rh->d3[b][rS_(x, j + 1)];
  return 0;
}
__device__ __forceinline__ int rMlmm_(const Ctx& x, int i, int j) { return x.rh->mmM[rBP_(x, j, i)][rS_(x, j - 1)][rS_(x, i + 1)]; }
#define hl_energy(i, j) rHl_(_x, (i), (j))
#define bl_energy(i, k, l, j, xr) rBl_(_x, (i), (k), (l), (j))
#define br_energy(i, k, l, j, xl) rBr_(_x, (i), (k), (l), (j))
#define il_energy(i, k, l, j) rIl_(_x, (i), (k), (l), (j))
#define sr_energy(i, j) rSr_(_x, (i), (j))
#define termau_energy(i, j) rTermau_(_x, (i), (j))
#define ext_mismatch_energy(i, j) rExt_(_x, (i), (j))
#define ml_mismatch_energy(i, j) rMlmm_(_x, (i), (j))
#define ml_energy() (_x.rh->ml_closing)
#define ul_energy() (_x.rh->ml_intern)
)";
}

bool generate(Model& m, std::string& src, std::string& err) {
  std::vector<std::vector<AltCode>> alts(m.tabs.size());
  std::vector<std::map<std::string, std::string>> pre(m.tabs.size());
  std::vector<std::vector<std::array<std::string, 4>>> pre_decl(m.tabs.size());
  // second pass: every reference fixed, so the reads are hoisted
  for (int pass = 0; pass < 2; ++pass)
  for (int t : m.order) {
    if (pass == 1 && !m.fixed_refs) break;
    alts[t].clear();
    const Tab& T = m.tabs[t];
    std::vector<NP> flat;
    if (!flatten(m, T.p, true, flat, err)) return false;
    if ((int)flat.size() != T.alt) { err = "internal: alternative count"; return false; }
    for (int q = 0; q < (int)flat.size(); ++q) {
      Gen g(m, true);
      if (pass == 1) {
        g.pre = &pre[t];
        g.pre_decl = &pre_decl[t];
      }
      AltCode o;
      o.rule = T.id + q;
      std::vector<std::string> vals;
      if (!g.gen(flat[q].get(), "i", "j", o, vals)) { err = g.err; return false; }
      if (!g.fixed) m.fixed_refs = false;
      if (vals.size() != 1) {
        err = "alternative " + std::to_string(o.rule) + " of '" + T.name + "' does not yield one answer (map the concatenation with a function)";
        return false;
      }
      std::function<bool(const AltCode&)> any_loops = [&](const AltCode& x) {
        if (!x.loops.empty()) return true;
        for (auto& N : x.nested)
          for (auto& a : N.alts)
            if (any_loops(*a)) return true;
        return false;
      };
      if (any_loops(o)) m.loops = true;
      o.value = vals[0];
      alts[t].push_back(o);
    }
  }
  // one track: the alternatives whose split candidates strictly inside a
  // tile's row and column blocks can be taken as a middle product (one split
  // loop from i + small to j - small, every reference through the split
  // variable); additive answers only when any order gives the same sum
  std::vector<std::vector<AltCode>> malts(m.tabs.size());
  std::vector<std::set<int>> mrule(m.tabs.size());
  std::vector<Gen::MRef> mrefs;
  std::set<int> minputs;
  bool has_mid = false;
  int mid_lo = 0, mid_hi = 0;  // the middle product's splits: [i0 + 32 + mid_lo, j0 - 1 - mid_hi]
  for (int t : m.order) {
    if (m.tracks != 1 || !m.any_order()) break;
    const Tab& T = m.tabs[t];
    std::vector<NP> flat;
    if (!flatten(m, T.p, true, flat, err)) return false;
    for (int q = 0; q < (int)flat.size(); ++q) {
      for (int final_pass = 0; final_pass < 2; ++final_pass) {
        std::vector<Gen::MRef> scratch;
        std::set<int> scratch_in;
        Gen g(m, false);
        g.mrefs = final_pass ? &mrefs : &scratch;
        g.minputs = final_pass ? &minputs : &scratch_in;
        AltCode o;
        o.rule = T.id + q;
        std::vector<std::string> vals;
        if (!g.gen(flat[q].get(), "i", "j", o, vals) || vals.size() != 1) break;
        if (!g.mid_ok || o.loops.size() != 1 || (!final_pass && scratch.empty())) break;
        const Loop& L = o.loops[0];
        int cl = 0, ch = 0;
        if (L.lo.size() != 1 || L.hi.size() != 1 || !lin_is(L.lo[0], "i", cl) || !lin_is(L.hi[0], "j", ch) ||
            cl > Gen::MIDM + 1 || ch < -(Gen::MIDM + 1))
          break;
        if (final_pass) {
          o.value = vals[0];
          malts[t].push_back(o);
          mrule[t].insert(o.rule);
          has_mid = true;
          // every cell of the tile has the whole range inside its loop bounds
          mid_lo = std::max({mid_lo, g.need_lo, cl - 1});
          mid_hi = std::max({mid_hi, g.need_hi, -ch - 1});
        }
      }
    }
  }
  // Left terms of the middle product: a call whose arguments are a staged
  // left operand (i, k) and other terms of i and the split alone (no j, no
  // right operand) has one value per row and split, shared by the row's 32
  // cells and by every rule that makes it; it is computed once per staged
  // block into LDS (mU<h>) when its type is the answer type, and read there
  std::vector<MidHoist> mhoist;
  for (int t : m.order)
    for (auto& o : malts[t]) o.value = hoist_left(o.value, o.loops[0].var, mhoist);
  size_t mid_lds = 0;
  {
    mid_lds += mhoist.size() * 1024 * (size_t)m.answer_bytes;
    for (auto& r : mrefs) mid_lds += 1024 * (size_t)m.answer_bytes + (r.rules ? 2048 : 0);
    mid_lds += minputs.size() * (32 + 2 * Gen::MIDM) * (size_t)m.alphabet_bytes;
    if (mid_lds > 48 * 1024) {  // too many operands to stage: the wavefront takes every split
      for (auto& v : malts) v.clear();
      for (auto& v : mrule) v.clear();
      mrefs.clear();
      minputs.clear();
      mhoist.clear();
      has_mid = false;
      mid_lds = 0;
    }
  }
  // the in-tile form of every alternative (one track), reading other cells
  // through the LDS region of the tile: rows i0 .. i0 + 31 + H, columns
  // j0 - H .. j0 + 31 -- the tile and, for H = 32, its neighbours (I, J - 1),
  // (I + 1, J - 1), (I + 1, J), all finished before the tile starts
  std::vector<std::vector<AltCode>> alts_tile(m.tabs.size());
  std::set<int> tile_rules;
  std::set<int> strip_l, strip_r;  // tabulations with a left / right strip in LDS
  size_t wcap_big = 16384;  // bytes of the whole-input copy of a 1024-thread tile
  int reg_h = -1;  // H, or -1: no region (every read from the tables in memory)
  std::set<int> mid_rule_ids;
  for (auto& v : mrule) mid_rule_ids.insert(v.begin(), v.end());
  // the in-tile alternatives: a first pass finds which tabulations' rule
  // entries the region needs; after the region and strips are sized, the
  // second pass generates them with the direct LDS reads
  std::vector<std::map<std::string, std::string>> pre_tile(m.tabs.size());
  std::vector<std::vector<std::array<std::string, 4>>> pre_decl_tile(m.tabs.size());
  auto gen_tile = [&](int reg, bool fast) {
    for (auto& v : alts_tile) v.clear();
    for (auto& v : pre_tile) v.clear();
    for (auto& v : pre_decl_tile) v.clear();
    for (int t : m.order) {
      const Tab& T = m.tabs[t];
      std::vector<NP> flat;
      if (!flatten(m, T.p, true, flat, err)) return false;
      for (int q = 0; q < (int)flat.size(); ++q) {
        Gen g(m, true);
        g.tile = true;
        g.tile_rules = &tile_rules;
        if (fast) {
          g.pre = &pre_tile[t];
          g.pre_decl = &pre_decl_tile[t];
          g.t_reg_h = reg;
          g.t_sl = &strip_l;
          g.t_sr = &strip_r;
          g.t_mid = &mid_rule_ids;
          g.t_ml = mid_lo;
          g.t_mh = mid_hi;
        }
        AltCode o;
        o.rule = T.id + q;
        std::vector<std::string> vals;
        if (!g.gen(flat[q].get(), "i", "j", o, vals) || vals.size() != 1) { err = g.err; return false; }
        o.value = vals[0];
        alts_tile[t].push_back(o);
      }
    }
    return true;
  };
  if (m.tracks == 1) {
    if (!gen_tile(-1, false)) return false;
    size_t per = 0, core = 0;  // region bytes per cell; the tile's rule and split arrays
    for (int t : m.order) {
      per += (size_t)m.answer_bytes + (tile_rules.count(t) ? 2 : 0);
      core += 1024 * (2 + 4 * (size_t)m.tabs[t].cat);
    }
    mid_lds += core + 256 * (size_t)m.alphabet_bytes;  // and the input windows
    // strips for the operands of the middle products: their splits left
    // to the in-tile pass read tiles (I, I), (I, I + 1) (left operands) and
    // (J - 1, J), (J, J) (right operands)
    size_t strips = 0;
    for (auto& r : mrefs) (r.left ? strip_l : strip_r).insert(r.t);
    for (int t : strip_l) strips += 2048 * ((size_t)m.answer_bytes + (tile_rules.count(t) ? 2 : 0));
    for (int t : strip_r) strips += 2048 * ((size_t)m.answer_bytes + (tile_rules.count(t) ? 2 : 0));
    // a workgroup may hold up to 160 KB of LDS; above 80 KB one workgroup
    // runs per CU, which the 1024-thread tiles need anyway, so a large
    // grammar (several tabulations, energy tables) takes the whole region
    // and one tile per CU rather than reading its halo from memory
    // The whole-input copy of a 1024-thread tile is 16 KB, or 8 KB where
    // that keeps the strips of a large grammar in LDS
    size_t big = 64 * 64 * per + mid_lds + 16384;
    const size_t head = m.rna ? offsetof(DevRna, int21) : 0;  // the energy tables' head staged per workgroup
    if (big + strips > 128 * 1024 && big > 96 * 1024 && big + strips + head + 1024 > 160 * 1024 &&
        big - 8192 + strips + head + 1024 <= 160 * 1024) {
      big -= 8192;
      wcap_big = 8192;
    }
    if (big + strips <= 128 * 1024) reg_h = 32;
    else if (big <= 96 * 1024) reg_h = 32, strip_l.clear(), strip_r.clear();
    else if (big + strips + head + 1024 <= 160 * 1024) reg_h = 32;
    else {
      strip_l.clear();
      strip_r.clear();
      if (big + head + 1024 <= 160 * 1024) reg_h = 32;
      else if (32 * 32 * per + mid_lds + wcap_big <= 96 * 1024) reg_h = 0;
    }
    if (!gen_tile(reg_h, true)) return false;
  }
  // third pass: the register form of the band kernel, when every reference
  // allows it
  std::vector<std::vector<AltCode>> alts_r(m.tabs.size());
  std::vector<std::map<std::string, std::string>> pre_r(m.tabs.size());
  std::vector<std::vector<std::array<std::string, 4>>> pre_decl_r(m.tabs.size());
  std::set<std::array<int, 3>> rset;
  bool reg_band = m.tracks == 2 && !m.loops && m.fixed_refs;
  for (int t : m.order) {
    if (!reg_band) break;
    const Tab& T = m.tabs[t];
    std::vector<NP> flat;
    if (!flatten(m, T.p, true, flat, err)) return false;
    for (int q = 0; q < (int)flat.size(); ++q) {
      Gen g(m, true);
      g.pre = &pre_r[t];
      g.pre_decl = &pre_decl_r[t];
      g.regs = &rset;
      AltCode o;
      o.rule = T.id + q;
      std::vector<std::string> vals;
      if (!g.gen(flat[q].get(), "i", "j", o, vals) || vals.size() != 1) { err = g.err; return false; }
      if (!g.reg_ok) reg_band = false;
      o.value = vals[0];
      alts_r[t].push_back(o);
    }
  }
  m.reg_band = reg_band;
  const int NT = (int)m.tabs.size();
  std::ostringstream s;
  s << "// generated by libdpx from an ADP grammar (hipRTC provides the HIP runtime declarations)\n"
       "#define GLOBAL_ __attribute__((address_space(1)))\n"
       "struct T2ii { int _1, _2; };\n"
    << m.header << "\n"
    << "typedef " << m.alphabet << " input_t;\n"
    << "typedef " << m.answer << " answer_t;\n"
    << "static_assert(sizeof(input_t) == " << m.alphabet_bytes << ", \"alphabet size\");\n"
    << "static_assert(sizeof(answer_t) == " << m.answer_bytes << ", \"answer size\");\n"
    << "struct Dims { int n1, n2; };\n"
    << "struct Tabs { void* c[" << NT << "]; short* r[" << NT << "]; int* p[" << NT << "][2]; };\n"
    << "struct Rec { int i, j, rule, pos[2]; };\n"
    << "// one instance of a batch: inputs, outputs and tables (host writes this layout)\n"
    << "struct Inst { const input_t* in1; const input_t* in2; answer_t* ans; int* out; Rec* tr; const void* rna; "
       "const int* hl; Dims z; int cap, pad; Tabs T; };\n"
    << "static_assert(__builtin_offsetof(Inst, z) == 56 && __builtin_offsetof(Inst, cap) == 64 && "
       "__builtin_offsetof(Inst, T) == 72 && sizeof(Inst) == 72 + " << 32 * NT << ", \"instance layout\");\n"
    << rna_prelude(m)
    << "// what every algebra function sees: the inputs, their sizes, the energy tables\n"
       "#define LDS_ __attribute__((address_space(3)))\n"
       "struct Ctx { const input_t* in1; const input_t* in2; int n1, n2; const RnaPar* rna; const int* hl; int wa, wb, wn;\n"
       "             const LDS_ input_t* wp; const LDS_ RnaHead* rh; };\n"
    << (m.tracks == 1 ? "#define M_W (_x.n1 + 1)\n#define M_H (_x.n1 + 1)\n" : "#define M_H (_x.n1 + 1)\n#define M_W (_x.n2 + 1)\n")
    // one track: input reads go through a view that serves a tile's two
    // input windows from LDS (origins wa, wb, stored + 2^30 so that a Ctx
    // built without them, wa = wb = 0, never matches) and the rest from memory
    << (m.tracks == 1
            ? "__shared__ input_t tinA_[128], tinB_[128];\n"

              "// the loads that miss the LDS copies sit behind a call, so the code\n"
              "// around an LDS read never waits for memory loads it does not use\n"
              "template <class V> __device__ __attribute__((noinline, pure)) V gld_(const V* p, long long x) {\n"
              "  return ((const GLOBAL_ V*)p)[x];\n}\n"
              "struct InV {\n"
              "  const input_t* g; int wa, wb, wn; const LDS_ input_t* wp;  // wp: the whole input, when it fits\n"
              "  __device__ __forceinline__ input_t operator[](int x) const {\n"
              "    if (wn) return wp[min(max(x, 0), wn - 1)];\n"
              "    const unsigned a = (unsigned)(x - wa + (1 << 30));\n"
              "    if (a < 128u) return tinA_[a];\n"
              "    const unsigned b = (unsigned)(x - wb + (1 << 30));\n"
              "    if (b < 128u) return tinB_[b];\n"
              "    return gld_(g, x);\n"
              "  }\n"
              "};\n"
              "#define IN1V_(x) (InV{(x).in1, (x).wa, (x).wb, (x).wn, (x).wp})\n"
              "#define IN1T_ const InV\n"
            : "#define IN1V_(x) ((x).in1)\n#define IN1T_ const input_t*\n")
    << (m.rna ? rna_functions() : std::string());
  if (m.tracks == 1)
    // 32 x 32 tiles, each a contiguous row-major record (a tile row is one
    // 128-byte line for 4-byte answers), tiles (I, J) numbered by tile
    // diagonal J - I, then I; windowed grammars keep the tile diagonals that
    // hold a cell of the window
    s << "#define TB 32\n"
         "__device__ __forceinline__ long long tile_ix_(int nb, int I, int J) {\n"
         "  const long long D = J - I;\n"
         "  return D * nb - D * (D - 1) / 2 + I;\n}\n"
         "__device__ __forceinline__ long long idx_(const Dims& z, int i, int j) {\n"
         "  const int nb = (z.n1 + TB) >> 5;\n"
         "  return (tile_ix_(nb, i >> 5, j >> 5) << 10) + ((i & 31) << 5) + (j & 31);\n}\n";
  else
    // bands of 64 rows, each stored by anti-diagonal (r + j), lanes (rows)
    // contiguous: one step of the band pipeline writes 64 adjacent cells, and
    // so does one anti-diagonal of a band in the diagonal kernels
    s << "__device__ __forceinline__ long long idx_(const Dims& z, int i, int j) {\n"
         "  const int b = i >> 6, r = i & 63;\n"
         "  return ((long long)b * (z.n2 + 64) + (r + j)) * 64 + r;\n}\n";
  // (i, j) when it is a cell of the tables, else cell 0: in bounds for any i, j
  if (m.tracks == 1)
    s << "__device__ __forceinline__ long long sidx_(const Dims& z, int i, int j) {\n"
      << "  const int D = " << (m.window > 0 ? "min(" + std::to_string(m.window) + ", z.n1)" : std::string("z.n1")) << ";\n"
      << "  return i >= 0 && i <= j && j <= z.n1 && j - i <= D ? idx_(z, i, j) : 0;\n}\n";
  else
    s << "__device__ __forceinline__ long long sidx_(const Dims& z, int i, int j) {\n"
         "  return i >= 0 && i <= z.n1 && j >= 0 && j <= z.n2 ? idx_(z, i, j) : 0;\n}\n";
  for (auto& f : m.fns) {
    if (!f.args.empty()) {
      s << "template <";
      for (size_t q = 0; q < f.args.size(); ++q) s << (q ? ", " : "") << "class A" << q;
      s << ">\n";
    }
    s << "__device__ __forceinline__ auto f_" << f.name << "(const Ctx& _x";
    for (size_t q = 0; q < f.args.size(); ++q) s << ", A" << q << " " << f.args[q];
    s << ") {\n  IN1T_ _in1 = IN1V_(_x);\n  const input_t* _in2 = _x.in2;\n  (void)_in1; (void)_in2;\n  " << f.body << "\n}\n";
  }
  auto better = [&](const std::string& c, const std::string& c0) {
    std::string cmp;
    switch (m.h.k) {
      case A_MAX: cmp = c + ">" + c0; break;
      case A_MIN: cmp = c + "<" + c0; break;
      case A_MAXBY: cmp = "f_" + m.fns[m.h.fn].name + "(_x," + c + ")>f_" + m.fns[m.h.fn].name + "(_x," + c0 + ")"; break;
      case A_MINBY: cmp = "f_" + m.fns[m.h.fn].name + "(_x," + c + ")<f_" + m.fns[m.h.fn].name + "(_x," + c0 + ")"; break;
      case A_SUM: case A_COUNT: cmp = "false"; break;
    }
    return cmp;
  };
  // A cell is evaluated by a group of G lanes of one wave (G = 1, 8 or 64,
  // chosen per diagonal): the outer split loop of each rule is strided over
  // the group, rules without a loop run on lane 0, and the group's
  // candidates are merged by value, then by (rule, outer split, inner split)
  // -- the first strictly better candidate of the sequential loop order.
  s << "template <class V> __device__ __forceinline__ V shfl_up1_(V v) {  // lane l gets lane l - 1's v\n"
       "  constexpr int N = (sizeof(V) + 3) / 4;\n"
       "  int u[N] = {};\n"
       "  __builtin_memcpy(u, &v, sizeof(V));\n"
       "  for (int q = 0; q < N; ++q) u[q] = __builtin_amdgcn_update_dpp(0, u[q], 0x138, 0xf, 0xf, false);  // wave_shr:1\n"
       "  __builtin_memcpy(&v, u, sizeof(V));\n"
       "  return v;\n}\n"
       "template <class V> __device__ __forceinline__ V shfl_idx_(V v, int l) {  // lane l's v\n"
       "  constexpr int N = (sizeof(V) + 3) / 4;\n"
       "  int u[N] = {};\n"
       "  __builtin_memcpy(u, &v, sizeof(V));\n"
       "  for (int q = 0; q < N; ++q) u[q] = __shfl(u[q], l, 64);\n"
       "  __builtin_memcpy(&v, u, sizeof(V));\n"
       "  return v;\n}\n"
       "template <class V> __device__ __forceinline__ V readlane_(V v, int l) {  // l is wave-uniform\n"
       "  constexpr int N = (sizeof(V) + 3) / 4;\n"
       "  int u[N] = {};\n"
       "  __builtin_memcpy(u, &v, sizeof(V));\n"
       "  for (int q = 0; q < N; ++q) u[q] = __builtin_amdgcn_readlane(u[q], l);\n"
       "  __builtin_memcpy(&v, u, sizeof(V));\n"
       "  return v;\n}\n";
  s << "template <class V> __device__ __forceinline__ void pin_(const V& v) {\n"
       "  int u[(sizeof(V) + 3) / 4] = {};\n"
       "  __builtin_memcpy(u, &v, sizeof(V));\n"
       "  for (int q = 0; q < (int)((sizeof(V) + 3) / 4); ++q) __asm__ volatile(\"\" ::\"v\"(u[q]));\n}\n";
  s << "template <int CTRL, class V> __device__ __forceinline__ V dpp_(V v) {\n"
       "  constexpr int N = (sizeof(V) + 3) / 4;\n"
       "  int u[N] = {};\n"
       "  __builtin_memcpy(u, &v, sizeof(V));\n"
       "  for (int q = 0; q < N; ++q) u[q] = __builtin_amdgcn_update_dpp(0, u[q], CTRL, 0xf, 0xf, false);\n"
       "  __builtin_memcpy(&v, u, sizeof(V));\n"
       "  return v;\n}\n"
       "// level lv of a group all-reduce: lane l's partner l ^ 1, l ^ 2, then\n"
       "// 7 - l within 8 lanes and 15 - l within 16 lanes\n"
       "template <class V> __device__ __forceinline__ V lvmove_(V v, int lv) {\n"
       "  switch (lv) {\n"
       "    case 0: return dpp_<0xB1>(v);   // quad_perm [1,0,3,2]\n"
       "    case 1: return dpp_<0x4E>(v);   // quad_perm [2,3,0,1]\n"
       "    case 2: return dpp_<0x141>(v);  // row_half_mirror\n"
       "    default: return dpp_<0x140>(v); // row_mirror\n"
       "  }\n}\n";
  s << "template <class V> __device__ __forceinline__ V shfl_xor_(V v, int m, int w) {\n"
       "  constexpr int N = (sizeof(V) + 3) / 4;\n"
       "  int u[N] = {};\n"
       "  __builtin_memcpy(u, &v, sizeof(V));\n"
       "  for (int q = 0; q < N; ++q) u[q] = __shfl_xor(u[q], m, w);\n"
       "  __builtin_memcpy(&v, u, sizeof(V));\n"
       "  return v;\n}\n";
  int used_tabs = 0;
  for (auto& T : m.tabs) used_tabs += T.used ? 1 : 0;
  // the cell's results in registers: read by the later tabulations of the
  // cell, stored by the last one (DEFER) or each by its own
  // table and input accesses through global pointers (not flat): they count
  // only in vmcnt, so the lane shifts (LDS counter) never wait for them

  s << "struct CellV {";
  for (int t : m.order) {
    s << " answer_t c" << t << "; short r" << t << ";";
    for (int q = 0; q < m.tabs[t].cat; ++q) s << " int p" << t << "_" << q << ";";
  }
  s << " };\n"
       "__device__ __forceinline__ void store_at(const Tabs& T, long long x, const CellV& cv) {\n";
  for (int t : m.order) {
    s << "  ((GLOBAL_ answer_t*)T.c[" << t << "])[x] = cv.c" << t << ";\n  ((GLOBAL_ short*)T.r[" << t << "])[x] = cv.r" << t << ";\n";
    for (int q = 0; q < m.tabs[t].cat; ++q)
      s << "  ((GLOBAL_ int*)T.p[" << t << "][" << q << "])[x] = cv.p" << t << "_" << q << ";\n";
  }
  s << "}\n"
       "__device__ __forceinline__ void store_cell(const Dims& z, const Tabs& T, int i, int j, const CellV& cv) {\n"
       "  store_at(T, idx_(z, i, j), cv);\n}\n";
  // a nested aggregation's block: its accumulator, its alternative's loops
  // (one lane, in order) and conditions by loop level, the first strictly
  // better candidate kept (the Scala aggregation's first maximum)
  auto better_of = [&](const Agg& h, const std::string& c, const std::string& c0) {
    switch (h.k) {
      case A_MAX: return c + ">" + c0;
      case A_MIN: return c + "<" + c0;
      case A_MAXBY: return "f_" + m.fns[h.fn].name + "(_x," + c + ")>f_" + m.fns[h.fn].name + "(_x," + c0 + ")";
      case A_MINBY: return "f_" + m.fns[h.fn].name + "(_x," + c + ")<f_" + m.fns[h.fn].name + "(_x," + c0 + ")";
      default: return std::string("false");
    }
  };
  std::function<void(const AltCode::Nest&, const std::string&)> emit_nest = [&](const AltCode::Nest& N, const std::string& ind0) {
    const std::string C = "_g" + N.id + "c", R = "_g" + N.id + "r", W = "_g" + N.id + "w";
    size_t npos = 0;
    for (auto& a : N.alts) npos = std::max(npos, a->pos.size());
    s << ind0 << "answer_t " << C << "{};\n" << ind0 << "int " << R << " = -1, " << W << " = -1;\n";
    for (size_t q = 0; q < npos; ++q) s << ind0 << "int _g" << N.id << "p" << q << " = 0;\n";
    for (int w = 0; w < (int)N.alts.size(); ++w) {
    const AltCode& in = *N.alts[w];
    std::string ind = ind0;
    s << ind << "{\n";
    int close = 1;
    const int NL = (int)in.loops.size();
    std::vector<std::vector<std::string>> at(NL + 1);
    auto level = [&](const std::string& c) {
      int lv = 0;
      for (int q = 0; q < NL; ++q) {
        const std::string& v = in.loops[q].var;
        for (size_t pos = c.find(v); pos != std::string::npos; pos = c.find(v, pos + 1)) {
          const char nx = pos + v.size() < c.size() ? c[pos + v.size()] : ' ';
          const char pv = pos ? c[pos - 1] : ' ';
          if (!isalnum((unsigned char)nx) && nx != '_' && !isalnum((unsigned char)pv) && pv != '_') { lv = q + 1; break; }
        }
      }
      return lv;
    };
    for (auto& c : in.arith) at[level(c)].push_back(c);
    for (size_t q = 0; q < in.other.size(); ++q) at[in.other_free[q] ? level(in.other[q]) : NL].push_back(in.other[q]);
    auto open_if = [&](int lv) {
      if (at[lv].empty()) return;
      ind += "  ";
      s << ind << "if (" << join(at[lv], " && ") << ") {\n";
      ++close;
    };
    open_if(0);
    for (int q = 0; q < NL; ++q) {
      const Loop& L = in.loops[q];
      ind += "  ";
      s << ind << "for (int " << L.var << " = " << bound(L.lo, "max") << ", " << L.var << "_u = " << bound(L.hi, "min") << "; "
        << L.var << " <= " << L.var << "_u; ++" << L.var << ") {\n";
      ++close;
      open_if(q + 1);
    }
    ind += "  ";
    std::vector<std::string> post;
    for (auto& M : in.nested) {
      emit_nest(M, ind);
      post.push_back("(_g" + M.id + "w == " + std::to_string(M.sib) + ")");
    }
    if (!post.empty()) {
      s << ind << "if (" << join(post, " && ") << ") {\n";
      ++close;
      ind += "  ";
    }
    s << ind << "const answer_t _c = " << in.value << ";\n"
      << ind << "if (" << better_of(N.h, "_c", C) << " || " << R << " == -1) { " << C << " = _c; " << R << " = 1; " << W
      << " = " << w << ";";
    for (size_t q = 0; q < in.pos.size(); ++q) s << " _g" << N.id << "p" << q << " = " << in.pos[q] << ";";
    s << " }\n";
    for (; close > 0; --close) {
      ind.resize(ind.size() - 2);
      s << ind << "}\n";
    }
    }
  };
  auto emit_tabs = [&](const std::vector<std::vector<AltCode>>& A, const std::vector<std::vector<std::array<std::string, 4>>>& PD,
                       const std::string& sx, const std::string& extra, bool tile) {
  for (int t : m.order) {
    const Tab& T = m.tabs[t];
    // tile: the in-tile form of a one-track cell -- alternatives with a
    // middle product skip the split points [_mlo, _mhi] (the tile's middle
    // product took them) and the cell's middle-product partial, left in its
    // table entry, joins the lanes' merge
    if (!tile) {
    s << "// ---- " << T.name << "[i,j]: rules " << T.id << ".." << T.id + T.alt - 1 << "\n"
      << "struct Pre" << sx << "_" << t << " {";
    for (auto& d : PD[t]) s << " " << d[0] << " " << d[1] << ";";
    s << " int pad_; };\n"
      << "__device__ __forceinline__ void load" << sx << "_" << t << "(const Ctx& _x, const Dims& z, const Tabs& T, int i, int j, Pre" << sx << "_" << t
      << "& p) {\n  const GLOBAL_ input_t* _in1 = (const GLOBAL_ input_t*)_x.in1;\n"
         "  const GLOBAL_ input_t* _in2 = (const GLOBAL_ input_t*)_x.in2;\n  (void)_in1; (void)_in2;\n";
    for (auto& d : PD[t]) s << "  p." << d[1] << " = " << d[2] << ";\n";
    s << "  p.pad_ = 0;\n}\n"
      << "__device__ __forceinline__ void pin" << sx << "_" << t << "(const Pre" << sx << "_" << t << "& p) {";
    for (auto& d : PD[t]) s << " pin_(p." << d[1] << ");";
    s << " (void)p; }\n";
    }
    s << "template <int G, bool DEFER> __device__ __forceinline__ void tab" << (tile ? "T" : sx) << "_" << t
      << "(const Ctx& _x, const Dims& z, const Tabs& T, int i, int j, int lane, CellV& cv"
      << (tile ? std::string() : ", const Pre" + sx + "_" + std::to_string(t) + "& _p") << extra
      << ") {\n"
      << "  IN1T_ _in1 = IN1V_(_x);\n  const input_t* _in2 = _x.in2;\n  (void)_in1; (void)_in2;\n"
      << "  answer_t _c0" << (m.empty.empty() ? "{}" : " = (answer_t)(" + m.empty + ")") << ";\n"
      << "  int _r0 = -1, _p0 = 0, _p1 = 0, _o0 = 0, _o1 = 0;\n";
    if (tile)  // the cell's fixed-offset reads, issued together
      for (auto& d : PD[t]) s << "  const " << d[0] << " " << d[1] << " = " << d[2] << ";\n";
    else
      for (auto& d : PD[t]) s << "  const " << d[0] << " " << d[1] << " = _p." << d[1] << ";\n";
    if (!tile) s << "  (void)_p;\n";
    if (tile && !mrule[t].empty())  // the middle product's partial, loaded ahead of the rules
    {
      const std::string ts = std::to_string(t);
      s << "  answer_t _mq_c{};\n  int _mq_r = -1, _mq_p0 = 0, _mq_p1 = 0;\n"
           "  if (_mid) {\n";
      if (reg_h >= 0)
        s << "    const int e_ = (i - _ri0) * 32 + (j - _rj0 - RHALO_);\n"
          << "    _mq_c = tv_" << ts << "[(i - _ri0) * RW_ + (j - _rj0)];\n"
          << "    _mq_r = tcr_" << ts << "[e_];\n"
          << (T.cat > 0 ? "    _mq_p0 = tcp_" + ts + "_0[e_];\n" : std::string())
          << (T.cat > 1 ? "    _mq_p1 = tcp_" + ts + "_1[e_];\n" : std::string());
      else
        s << "    const long long x = idx_(z, i, j);\n"
          << "    _mq_c = ((const GLOBAL_ answer_t*)T.c[" << ts << "])[x];\n"
          << "    _mq_r = ((const GLOBAL_ short*)T.r[" << ts << "])[x];\n"
          << (T.cat > 0 ? "    _mq_p0 = ((const GLOBAL_ int*)T.p[" + ts + "][0])[x];\n" : std::string())
          << (T.cat > 1 ? "    _mq_p1 = ((const GLOBAL_ int*)T.p[" + ts + "][1])[x];\n" : std::string());
      s << "  }\n";
    }
    int nl_rule = 0;
    for (auto& o : A[t]) {
      std::string ind = "  ";
      s << ind << "{  // rule " << o.rule << "\n";
      int close = 1;
      if (o.loops.empty()) {
        // a rule without a split loop on one lane of the group, a different
        // lane for each such rule (the group merge orders by rule and split,
        // so the lane holding a candidate does not matter)
        ind += "  ";
        s << ind << "if (lane == (" << nl_rule++ << " % G)) {\n";
        ++close;
      }
      // each condition at the outermost loop level that binds its names
      // (index conditions and the guarded table tests anywhere, filters of
      // the cell itself before the loops, other filters innermost); pure, so
      // only the work changes, as in the reference's emit (CodeGen.scala:162-184)
      const int NL = (int)o.loops.size();
      std::vector<std::vector<std::string>> at(NL + 1);
      auto level = [&](const std::string& c) {
        int lv = 0;
        for (int q = 0; q < NL; ++q) {
          const std::string& v = o.loops[q].var;
          for (size_t pos = c.find(v); pos != std::string::npos; pos = c.find(v, pos + 1)) {
            const char nx = pos + v.size() < c.size() ? c[pos + v.size()] : ' ';
            const char pv = pos ? c[pos - 1] : ' ';
            if (!isalnum((unsigned char)nx) && nx != '_' && !isalnum((unsigned char)pv) && pv != '_') { lv = q + 1; break; }
          }
        }
        return lv;
      };
      for (auto& c : o.arith) at[level(c)].push_back(c);
      for (size_t q = 0; q < o.other.size(); ++q) at[o.other_free[q] ? level(o.other[q]) : NL].push_back(o.other[q]);
      auto open_if = [&](int lv) {
        if (at[lv].empty()) return;
        ind += "  ";
        s << ind << "if (" << join(at[lv], " && ") << ") {\n";
        ++close;
      };
      open_if(0);
      const bool split_mid = tile && mrule[t].count(o.rule);
      for (int q = 0; q < NL; ++q) {
        const Loop& L = o.loops[q];
        ind += "  ";
        if (q == 0 && split_mid) {
          // two runs: below and above the middle product's split points
          s << ind << "for (int _sg = 0; _sg < 2; ++_sg) {\n";
          ++close;
          ind += "  ";
          s << ind << "const int _lo = " << bound(L.lo, "max") << ", _hi = " << bound(L.hi, "min") << ";\n"
            << "#pragma unroll 4\n"
            << ind << "for (int " << L.var << " = (_sg ? max(_lo, _mhi + 1) : _lo) + lane, " << L.var
            << "_u = _sg ? _hi : min(_hi, _mlo - 1); " << L.var << " <= " << L.var << "_u; " << L.var << " += G) {\n";
        } else if (q == 0)
          s << ind << "for (int " << L.var << " = " << bound(L.lo, "max") << " + lane, " << L.var << "_u = " << bound(L.hi, "min")
            << "; " << L.var << " <= " << L.var << "_u; " << L.var << " += G) {\n";
        else
          s << ind << "for (int " << L.var << " = " << bound(L.lo, "max") << ", " << L.var << "_u = " << bound(L.hi, "min")
            << "; " << L.var << " <= " << L.var << "_u; ++" << L.var << ") {\n";
        ++close;
        open_if(q + 1);
      }
      ind += "  ";
      if (!o.nested.empty()) {
        std::vector<std::string> post;
        for (auto& N : o.nested) {
          emit_nest(N, ind);
          post.push_back("(_g" + N.id + "w == " + std::to_string(N.sib) + ")");
        }
        s << ind << "if (" << join(post, " && ") << ") {\n";
        ++close;
        ind += "  ";
      }
      const std::string take = "_r0 = " + std::to_string(o.rule) + "; _p0 = " + (o.pos.size() > 0 ? o.pos[0] : "0") +
                               "; _p1 = " + (o.pos.size() > 1 ? o.pos[1] : "0") + "; _o0 = " +
                               (o.loops.size() > 0 ? o.loops[0].var : "0") + "; _o1 = " +
                               (o.loops.size() > 1 ? o.loops[1].var : "0") + ";";
      if (m.h.k == A_SUM)  // the first candidate keeps the back-pointer
        s << ind << "const answer_t _c = " << o.value << ";\n"
          << ind << "if (_r0 == -1) { _c0 = answer_t(answer_t(0) + _c); " << take << " } else _c0 = answer_t(_c0 + _c);\n";
      else if (m.h.k == A_COUNT)
        s << ind << "if (_r0 == -1) { _c0 = answer_t(1); " << take << " } else _c0 = answer_t(_c0 + 1);\n";
      else
        s << ind << "const answer_t _c = " << o.value << ";\n"
          << ind << "if (" << better("_c", "_c0") << (m.empty.empty() ? " || _r0 == -1" : "") << ") { _c0 = _c; " << take
          << " }\n";
      for (; close > 0; --close) {
        ind.resize(ind.size() - 2);
        s << ind << "}\n";
      }
    }
    // all-reduce over the group: DPP lane moves within 16-lane rows (xor 1,
  // xor 2, half-row and row mirrors pair disjoint halves at each level),
  // lane shuffles across rows
  // only what differs between the lanes moves: the second split when some
  // rule has one, the order keys when they are not the splits themselves
  // (one split loop per rule, its variable the first split)
  bool need_p1 = false, o_is_p = true;
  for (auto& o : A[t]) {
    if (o.pos.size() >= 2 || o.loops.size() >= 2 || !o.nested.empty()) need_p1 = true;
    if (!o.loops.empty() && (o.pos.empty() || o.var_of_pos[0] != o.loops[0].var)) o_is_p = false;
  }
  o_is_p = o_is_p && !need_p1;
  s << "  for (int lv = 0; (1 << lv) < G; ++lv) {\n"
       "    answer_t c;\n"
       "    int r, p0, p1, o0, o1;\n"
       "    if (lv < 4) {\n"
       "      c = lvmove_(_c0, lv); r = lvmove_(_r0, lv); p0 = lvmove_(_p0, lv);\n"
    << (need_p1 ? "      p1 = lvmove_(_p1, lv);\n" : "      p1 = 0;\n")
    << (o_is_p ? "      o0 = p0; o1 = p1;\n" : "      o0 = lvmove_(_o0, lv); o1 = lvmove_(_o1, lv);\n")
    << "    } else {\n"
       "      const int off = 1 << lv;\n"
       "      c = shfl_xor_(_c0, off, G); r = __shfl_xor(_r0, off, G); p0 = __shfl_xor(_p0, off, G);\n"
    << (need_p1 ? "      p1 = __shfl_xor(_p1, off, G);\n" : "      p1 = 0;\n")
    << (o_is_p ? "      o0 = p0; o1 = p1;\n" : "      o0 = __shfl_xor(_o0, off, G); o1 = __shfl_xor(_o1, off, G);\n")
    << "    }\n"
       "    bool take;\n"
         "    if (r == -1) take = false;\n"
         "    else if (_r0 == -1) take = true;\n";
    if (m.additive())  // both lanes hold candidates: add them up, keep the first back-pointer
      s << "    else {\n"
           "      _c0 = answer_t(_c0 + c);\n"
           "      if (r < _r0 || (r == _r0 && (o0 < _o0 || (o0 == _o0 && o1 < _o1)))) { _r0 = r; _p0 = p0; _p1 = p1; _o0 = o0; _o1 = o1; }\n"
           "      take = false;\n"
           "    }\n";
    else
      s << "    else if (" << better("c", "_c0") << ") take = true;\n"
        << "    else if (" << better("_c0", "c") << ") take = false;\n"
        << "    else take = r < _r0 || (r == _r0 && (o0 < _o0 || (o0 == _o0 && o1 < _o1)));\n";
    s <<
         "    if (take) { _c0 = c; _r0 = r; _p0 = p0; _p1 = p1; _o0 = o0; _o1 = o1; }\n"
         "  }\n";
    if (tile && !mrule[t].empty()) {
      // the middle product's partial (value, rule, splits) of this cell,
      // merged as one more lane; its outer split is the split variable
      s << "  if (_mid) {\n"
           "    const answer_t c = _mq_c;\n"
           "    const int r = _mq_r, p0 = _mq_p0, p1 = _mq_p1;\n"
           "    int o0 = 0;\n"
           "    switch (r) {\n";
      for (auto& o : malts[t])
        for (size_t q = 0; q < o.var_of_pos.size(); ++q)
          if (o.var_of_pos[q] == o.loops[0].var) s << "      case " << o.rule << ": o0 = p" << q << "; break;\n";
      s << "      default: break;\n"
           "    }\n"
           "    bool take;\n"
           "    if (r == -1) take = false;\n"
           "    else if (_r0 == -1) take = true;\n";
      if (m.additive())
        s << "    else {\n"
             "      _c0 = answer_t(_c0 + c);\n"
             "      take = r < _r0 || (r == _r0 && (o0 < _o0 || (o0 == _o0 && 0 < _o1)));\n"
             "      if (take) { _r0 = r; _p0 = p0; _p1 = p1; _o0 = o0; _o1 = 0; }\n"
             "      take = false;\n"
             "    }\n";
      else
        s << "    else if (" << better("c", "_c0") << ") take = true;\n"
          << "    else if (" << better("_c0", "c") << ") take = false;\n"
          << "    else take = r < _r0 || (r == _r0 && (o0 < _o0 || (o0 == _o0 && 0 < _o1)));\n";
      s << "    if (take) { _c0 = c; _r0 = r; _p0 = p0; _p1 = p1; _o0 = o0; _o1 = 0; }\n"
           "  }\n";
    }
    s << "  (void)_o0; (void)_o1;\n"
      << "  cv.c" << t << " = _c0;\n  cv.r" << t << " = (short)_r0;\n";
    for (int q = 0; q < T.cat; ++q) s << "  cv.p" << t << "_" << q << " = _p" << q << ";\n";
    s << "  if (!DEFER && lane == 0) {\n";
    if (tile && reg_h >= 0) {
      // the tile's own cells stay in LDS; the tile goes to memory at its end
      s << "    const int e_ = (i - _ri0) * RW_ + (j - _rj0), c_ = (i - _ri0) * 32 + (j - _rj0 - RHALO_);\n"
        << "    tv_" << t << "[e_] = _c0;\n"
        << "    tcr_" << t << "[c_] = (short)_r0;\n";
      if (tile_rules.count(t)) s << "    trr_" << t << "[e_] = (short)_r0;\n";
      for (int q = 0; q < T.cat; ++q) s << "    tcp_" << t << "_" << q << "[c_] = _p" << q << ";\n";
    } else {
      s << "    const long long x = idx_(z, i, j);\n"
        << "    ((GLOBAL_ answer_t*)T.c[" << t << "])[x] = _c0;\n"
        << "    ((GLOBAL_ short*)T.r[" << t << "])[x] = (short)_r0;\n";
      for (int q = 0; q < T.cat; ++q) s << "    ((GLOBAL_ int*)T.p[" << t << "][" << q << "])[x] = _p" << q << ";\n";
    }
    s << "  }\n";
    if (used_tabs > 1 && !m.fixed_refs)  // a later tabulation of this cell may read it from another lane of the wave
      s << "  if (G > 1) { __builtin_amdgcn_s_waitcnt(0x0F70); __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, \"workgroup\"); }\n";
    s << "}\n";
  }
  };
  emit_tabs(alts, pre_decl, "", "", false);
  if (m.tracks == 1) {
    if (reg_h >= 0)
      s << "#define RH_ " << 32 + reg_h << "\n#define RW_ " << 32 + reg_h << "\n#define RHALO_ " << reg_h << "\n";
    for (int t : m.order) {
      const std::string ts = std::to_string(t);
      const bool sl = strip_l.count(t), sr = strip_r.count(t), rl = tile_rules.count(t);
      // per tabulation one array: the region, then the left strip (rows
      // i0 .. i0 + 31, columns i0 .. i0 + 63), then the right strip (rows
      // j0 - 32 .. j0 + 31, columns j0 .. j0 + 31)
      if (reg_h >= 0) {
        // the tile's own rule and split entries (its values are in the region)
        s << "__shared__ short tcr_" << ts << "[1024];\n";
        for (int q = 0; q < m.tabs[t].cat; ++q) s << "__shared__ int tcp_" << ts << "_" << q << "[1024];\n";
        s << "#define SLO_" << ts << " (RH_ * RW_)\n"
          << "#define SRO_" << ts << " (RH_ * RW_ + " << (sl ? 2048 : 0) << ")\n"
          << "#define TLEN_" << ts << " (RH_ * RW_ + " << (sl ? 2048 : 0) + (sr ? 2048 : 0) << ")\n"
          << "__shared__ answer_t tv_" << ts << "[TLEN_" << ts << "];\n";
        if (rl) s << "__shared__ short trr_" << ts << "[TLEN_" << ts << "];\n";
      }
      for (int rule = 0; rule < 2; ++rule) {
        if (rule)
          s << "__device__ __forceinline__ short rr_" << ts << "(const Dims& z, const Tabs& T, int ri0, int rj0, int a, int b) {\n";
        else
          s << "__device__ __forceinline__ answer_t rv_" << ts << "(const Dims& z, const Tabs& T, int ri0, int rj0, int a, int b) {\n";
        const bool lds = !rule || rl;
        if (reg_h >= 0 && lds)
          s << "  const unsigned ra = a - ri0, rb = b - rj0;\n"
               "  if (ra < RH_ && rb < RW_) return " << (rule ? "trr_" : "tv_") << ts << "[ra * RW_ + rb];\n";
        if (sl && lds)
          s << "  if ((unsigned)(a - ri0) < 32u && (unsigned)(b - ri0) < 64u) return " << (rule ? "trr_" : "tv_") << ts
            << "[SLO_" << ts << " + (a - ri0) * 64 + (b - ri0)];\n";
        if (sr && lds)
          s << "  { const int j0 = rj0 + RHALO_;\n"
               "    if ((unsigned)(a - j0 + 32) < 64u && (unsigned)(b - j0) < 32u) return " << (rule ? "trr_" : "tv_") << ts
            << "[SRO_" << ts << " + (a - j0 + 32) * 32 + (b - j0)]; }\n";
        s << "  (void)ri0; (void)rj0;\n"
          << "  return gld_((const " << (rule ? "short" : "answer_t") << "*)T." << (rule ? "r" : "c") << "[" << ts
          << "], idx_(z, a, b));\n}\n";
      }
    }
    emit_tabs(alts_tile, pre_decl_tile, "", ", bool _mid, int _mlo, int _mhi, int _ri0, int _rj0", true);
  }
  if (m.reg_band) {
    // what the register band kernel's cells read: R.v<t>_<di>_<dj> (value)
    // and R.r<t>_<di>_<dj> (rule) of cell (i - di, j - dj)
    s << "struct Regs {";
    for (auto& k : rset)
      s << " answer_t v" << k[0] << "_" << k[1] << "_" << k[2] << "; short r" << k[0] << "_" << k[1] << "_" << k[2] << ";";
    s << " int pad_; };\n";
    emit_tabs(alts_r, pre_decl_r, "R", ", const Regs& R", false);
  }
  // every tabulation's reads of other cells first, all in flight together
  // (pinned here rather than sunk into the rules), then the tabulations
  auto cell_body = [&](const std::string& ind, const std::string& G, const std::string& lane, const std::string& defer) {
    s << ind << "CellV cv;\n";
    for (int t : m.order) s << ind << "cv.r" << t << " = -1;\n";
    for (int t : m.order) s << ind << "Pre_" << t << " p" << t << ";\n" << ind << "load_" << t << "(_x, z, T, i, j, p" << t << ");\n";
    for (int t : m.order) s << ind << "pin_" << t << "(p" << t << ");\n";
    for (int t : m.order)
      s << ind << "tab_" << t << "<" << G << ", " << defer << ">(_x, z, T, i, j, " << lane << ", cv, p" << t << ");\n";
  };
  // fill: one diagonal per launch (d0 == d1 - 1), or a run of small
  // diagonals by one workgroup with a barrier between them (two tracks;
  // one-track grammars run on the tile queue)
  if (m.tracks == 2) {
    s << "template <int G> __device__ __forceinline__ void fill_(const Inst* __restrict__ I, int d0, int d1) {\n"
         "  const Inst& q = I[blockIdx.y];  // instances sorted by size, largest first: blockIdx.y < active ones\n"
         "  const Dims z = q.z;\n  const Tabs T = q.T;  // in registers: the stores below cannot change it\n"
         "  const Ctx _x{q.in1, q.in2, z.n1, z.n2, (const RnaPar*)q.rna, q.hl};\n"
         "  const int tid = blockIdx.x * blockDim.x + threadIdx.x, lane = tid & (G - 1);\n"
         "  for (int d = d0; d < d1; ++d) {\n"
         "    const int ilo = d > z.n2 ? d - z.n2 : 0, ihi = d < z.n1 ? d : z.n1, cnt = ihi - ilo + 1;\n"
         "    for (int t = tid / G; t < cnt; t += gridDim.x * blockDim.x / G) {\n"
         "      const int i = ilo + t, j = d - i;\n";
    cell_body("      ", "G", "lane", "false");
    s << "    }\n"
         "    if (d + 1 < d1) __syncthreads();\n"
         "  }\n}\n";
    for (int G : {1, 8, 64})
      s << "extern \"C\" __global__ void adp_fill_" << G << "(const Inst* I, int d0, int d1) { fill_<" << G << ">(I, d0, d1); }\n";
  }
  if (m.tracks == 1) {
    const int MIDM = Gen::MIDM;
    const std::string init = m.empty.empty() ? "answer_t{}" : "(answer_t)(" + m.empty + ")";
    // staged operands of the middle product
    s << "#define MIDM " << MIDM << "\n"
         "struct MidSmB {";
    for (size_t x = 0; x < mrefs.size(); ++x) {
      const std::string side = mrefs[x].left ? "mL" : "mR";
      s << " answer_t " << side << x << "[1024];";
      if (mrefs[x].rules) s << " short " << side << "R" << x << "[1024];";
    }
    for (int tr : minputs) s << " input_t mI" << tr << "[" << 32 + 2 * MIDM << "];";
    s << " int pad_; };\n";
    // the left terms: kept in LDS (MUH<h>_) when their type is the answer type
    s << "template <class A, class B> struct same_t_ { static constexpr bool v = false; };\n"
         "template <class A> struct same_t_<A, A> { static constexpr bool v = true; };\n";
    const std::string mu_params = "(const Ctx& _x, const Dims& z, const Tabs& T, IN1T_& _in1, const input_t* _in2, int i, int ";
    for (size_t h = 0; h < mhoist.size(); ++h) {
      const std::string hs = std::to_string(h);
      s << "__host__ __device__ auto mu_t" << hs << "_" << mu_params << mhoist[h].var << ", int _mr, int _mkk, const MidSmB& sm) -> decltype("
        << mhoist[h].expr << ");\n"
        << "constexpr bool MUH" << hs << "_ = same_t_<decltype(mu_t" << hs << "_(*(const Ctx*)nullptr, *(const Dims*)nullptr, "
        << "*(const Tabs*)nullptr, *(IN1T_*)nullptr, nullptr, 0, 0, 0, 0, *(const MidSmB*)nullptr)), answer_t>::v;\n";
    }
    s << "struct MidSm : MidSmB {";
    for (size_t h = 0; h < mhoist.size(); ++h) s << " answer_t mU" << h << "[MUH" << h << "_ ? 1024 : 1];";
    s << " int pad2_; };\n";
    for (size_t h = 0; h < mhoist.size(); ++h) {
      const std::string hs = std::to_string(h);
      s << "template <bool H> __device__ __forceinline__ auto mu" << hs << "_" << mu_params << mhoist[h].var
        << ", int _mr, int _mkk, const MidSm& sm) {\n"
        << "  (void)_x; (void)z; (void)T; (void)_in1; (void)_in2; (void)i; (void)" << mhoist[h].var << ";\n"
        << "  if constexpr (H) return sm.mU" << hs << "[_mr * 32 + _mkk];\n"
        << "  else return " << mhoist[h].expr << ";\n}\n";
    }
    if (has_mid) {
      // Middle product of tile (i0, j0): for every cell and alternative with
      // a middle product, the best candidate over the split points
      // [KLO, KHI] -- blocks of 32 splits whose left operand rows i0.. and
      // right operand columns j0.. are staged in LDS by the whole workgroup;
      // each thread owns CPT cells of one column (a wave: two rows of 32
      // cells, so a left-operand read is a broadcast and a right-operand
      // read 32 adjacent words).  Within an alternative the splits run in
      // increasing order, so "strictly better" keeps the first; the
      // alternatives then merge in rule order, and the partial goes to the
      // cell's table entry for the in-tile pass.
      s << "template <int CPT> __device__ __forceinline__ void middle_(const Ctx& _x, const Dims& z, const Tabs& T, int D, int i0,\n"
           "                                                        int j0, int KLO, int KHI, MidSm& sm) {\n"
           "  IN1T_ _in1 = IN1V_(_x);\n  const input_t* _in2 = _x.in2;\n  (void)_in1; (void)_in2;\n"
           "  constexpr int RS = 32 / CPT, NT_ = 32 * RS;\n"
           "  const int tid = threadIdx.x, _mc = tid & 31, rb = tid >> 5;\n"
           "  bool ok[CPT];\n";
      for (int t : m.order)
        for (auto& o : malts[t]) s << "  answer_t bc" << o.rule << "[CPT];\n  int bk" << o.rule << "[CPT];\n";
      s << "#pragma unroll\n"
           "  for (int u = 0; u < CPT; ++u) {\n"
           "    const int i = i0 + rb + RS * u, j = j0 + _mc;\n"
           "    ok[u] = j <= z.n1 && j - i <= D;\n";
      for (int t : m.order)
        for (auto& o : malts[t]) s << "    bc" << o.rule << "[u] = " << init << ";\n    bk" << o.rule << "[u] = -1;\n";
      s << "  }\n";
      // the operands of block K travel in registers: block K + 1 is fetched
      // while block K is used
      for (size_t x = 0; x < mrefs.size(); ++x) {
        s << "  answer_t pf" << x << "[CPT];\n";
        if (mrefs[x].rules) s << "  short pfr" << x << "[CPT];\n";
      }
      for (int tr : minputs) s << "  input_t pfi" << tr << " = input_t{};\n";
      std::ostringstream fetch;
      fetch << "#pragma unroll\n"
               "    for (int w = 0; w < CPT; ++w) {\n"
               "      const int e = tid + w * NT_, hi_ = e >> 5, lo_ = e & 31;\n";
      for (size_t x = 0; x < mrefs.size(); ++x) {
        const auto& r = mrefs[x];
        const std::string t = std::to_string(r.t);
        const std::string ix = r.left ? "sidx_(z, i0 + hi_ + " + std::to_string(r.a) + ", (KF << 5) + lo_ + " + std::to_string(r.b) + ")"
                                      : "sidx_(z, (KF << 5) + hi_ + " + std::to_string(r.a) + ", j0 + lo_ + " + std::to_string(r.b) + ")";
        fetch << "      {\n        const long long a_ = " << ix << ";\n"
              << "        pf" << x << "[w] = ((const GLOBAL_ answer_t*)T.c[" << t << "])[a_];\n";
        if (r.rules) fetch << "        pfr" << x << "[w] = ((const GLOBAL_ short*)T.r[" << t << "])[a_];\n";
        fetch << "      }\n";
      }
      fetch << "    }\n";
      for (int tr : minputs)
        fetch << "    if (tid < " << 32 + 2 * MIDM << ") pfi" << tr << " = ((const GLOBAL_ input_t*)_x.in" << tr
              << ")[min(max((KF << 5) - MIDM + tid, 0), z.n1)];\n";
      s << "  {\n    const int KF = KLO >> 5;\n" << fetch.str() << "  }\n"
        << "  for (int K = KLO >> 5; K <= (KHI >> 5); ++K) {\n"
           "    __syncthreads();  // the last block's operands are read\n"
           "#pragma unroll\n"
           "    for (int w = 0; w < CPT; ++w) {\n"
           "      const int e = tid + w * NT_;\n";
      for (size_t x = 0; x < mrefs.size(); ++x) {
        const std::string side = mrefs[x].left ? "mL" : "mR";
        s << "      sm." << side << x << "[e] = pf" << x << "[w];\n";
        if (mrefs[x].rules) s << "      sm." << side << "R" << x << "[e] = pfr" << x << "[w];\n";
      }
      s << "    }\n";
      for (int tr : minputs) s << "    if (tid < " << 32 + 2 * MIDM << ") sm.mI" << tr << "[tid] = pfi" << tr << ";\n";
      s << "    if (K < (KHI >> 5)) {\n      const int KF = K + 1;\n" << fetch.str() << "    }\n";
    s << "    __syncthreads();\n";
      if (!mhoist.empty()) {
        s << "    if (";
        for (size_t h = 0; h < mhoist.size(); ++h) s << (h ? " || " : "") << "MUH" << h << "_";
        s << ") {\n"
             "#pragma unroll\n"
             "      for (int w = 0; w < CPT; ++w) {\n"
             "        const int e = tid + w * NT_, _mr = e >> 5, _mkk = e & 31, i = i0 + _mr, k_ = (K << 5) + _mkk;\n"
             "        (void)i;\n"
             "        if (k_ >= KLO && k_ <= KHI) {\n";
        for (size_t h = 0; h < mhoist.size(); ++h) {
          // evaluated only where its staged left operands are valid cells
          std::vector<std::string> g;
          const std::string& ex = mhoist[h].expr;
          for (size_t q = ex.find("sm.mL"); q != std::string::npos; q = ex.find("sm.mL", q + 1)) {
            size_t r = q + 5;
            while (r < ex.size() && isdigit((unsigned char)ex[r])) ++r;
            if (r == q + 5) continue;
            const int x = std::stoi(ex.substr(q + 5, r - q - 5));
            if (!m.empty.empty()) g.push_back("!(sm.mL" + std::to_string(x) + "[e]==(answer_t)(" + m.empty + "))");
            else if (mrefs[x].rules) g.push_back("sm.mLR" + std::to_string(x) + "[e]!=-1");
          }
          s << "          if constexpr (MUH" << h << "_) {\n"
            << "            const int " << mhoist[h].var << " = k_;\n"
            << "            if (" << (g.empty() ? std::string("true") : join(g, " && ")) << ") sm.mU" << h << "[e] = " << ex << ";\n"
            << "          }\n";
        }
        s << "        }\n"
             "      }\n"
             "      __syncthreads();\n"
             "    }\n";
      }
      s << "    const int klo = max(KLO, K << 5), khi = min(KHI, (K << 5) + 31);\n"
           "    for (int k = klo; k <= khi; ++k) {\n"
           "      const int _mkk = k - (K << 5);\n"
           "#pragma unroll\n"
           "      for (int u = 0; u < CPT; ++u) {\n"
           "        const int _mr = rb + RS * u, i = i0 + _mr, j = j0 + _mc;\n"
           "        (void)i; (void)j; (void)_mr;\n"
           "        if (ok[u]) {\n";
      for (int t : m.order)
        for (auto& o : malts[t]) {
          std::vector<std::string> cond = o.arith;
          for (auto& c : o.other) cond.push_back(c);
          const std::string R = std::to_string(o.rule);
          s << "          {  // rule " << o.rule << "\n"
            << "            const int " << o.loops[0].var << " = k;\n";
          if (!cond.empty()) s << "            if (" << join(cond, " && ") << ")\n";
          s << "            {\n";
          if (m.h.k == A_SUM)
            s << "              const answer_t _c = " << o.value << ";\n"
              << "              if (bk" << R << "[u] < 0) { bc" << R << "[u] = answer_t(answer_t(0) + _c); bk" << R << "[u] = k; }\n"
              << "              else bc" << R << "[u] = answer_t(bc" << R << "[u] + _c);\n";
          else if (m.h.k == A_COUNT)
            s << "              if (bk" << R << "[u] < 0) { bc" << R << "[u] = answer_t(1); bk" << R << "[u] = k; }\n"
              << "              else bc" << R << "[u] = answer_t(bc" << R << "[u] + 1);\n";
          else
            s << "              const answer_t _c = " << o.value << ";\n"
              << "              if (" << better("_c", "bc" + R + "[u]") << (m.empty.empty() ? " || bk" + R + "[u] < 0" : "")
              << ") { bc" << R << "[u] = _c; bk" << R << "[u] = k; }\n";
          s << "            }\n"
               "          }\n";
        }
      s << "        }\n"
           "      }\n"
           "    }\n"
           "  }\n"
           "#pragma unroll\n"
           "  for (int u = 0; u < CPT; ++u) {\n"
           "    const int i = i0 + rb + RS * u, j = j0 + _mc;\n"
           "    if (!ok[u]) continue;\n"
           "    const long long x = idx_(z, i, j);\n"
           "    (void)x;\n";
      for (int t : m.order) {
        if (malts[t].empty()) continue;
        const Tab& T = m.tabs[t];
        s << "    {\n      answer_t pc = " << init << ";\n      int pr = -1, pp0 = 0, pp1 = 0;\n";
        for (auto& o : malts[t]) {
          const std::string R = std::to_string(o.rule);
          const std::string setp = "const int " + o.loops[0].var + " = bk" + R + "[u]; pr = " + R + "; pp0 = " +
                                   (o.pos.size() > 0 ? o.pos[0] : "0") + "; pp1 = " + (o.pos.size() > 1 ? o.pos[1] : "0") + ";";
          s << "      if (bk" << R << "[u] >= 0) {\n";
          if (m.additive())
            s << "        if (pr == -1) { pc = bc" << R << "[u]; " << setp << " } else pc = answer_t(pc + bc" << R << "[u]);\n";
          else
            s << "        if (pr == -1 || " << better("bc" + R + "[u]", "pc") << ") { pc = bc" << R << "[u]; " << setp << " }\n";
          s << "      }\n";
        }
        if (reg_h >= 0) {
          s << "      tv_" << t << "[(i - i0) * RW_ + (j - j0 + RHALO_)] = pc;\n      tcr_" << t << "[(i - i0) * 32 + (j - j0)] = (short)pr;\n";
          for (int q = 0; q < T.cat; ++q) s << "      tcp_" << t << "_" << q << "[(i - i0) * 32 + (j - j0)] = pp" << q << ";\n";
        } else {
          s << "      ((GLOBAL_ answer_t*)T.c[" << t << "])[x] = pc;\n      ((GLOBAL_ short*)T.r[" << t << "])[x] = (short)pr;\n";
          for (int q = 0; q < T.cat; ++q) s << "      ((GLOBAL_ int*)T.p[" << t << "][" << q << "])[x] = pp" << q << ";\n";
        }
        s << "      (void)pp0; (void)pp1;\n    }\n";
      }
      s << "  }\n}\n";
    }
    // One persistent launch fills every tile of every instance.  Tiles wait
    // in one queue of 64-bit items (instance, I, J): the diagonal tiles are
    // seeded, and tile (I, J) is appended by whichever of (I, J - 1) and
    // (I + 1, J) finishes second -- between them they cover every cell a cell
    // of the tile reads outside the tile.  A workgroup claims the next slot
    // and waits (bounded) until it holds a tile; a tile's stores are
    // published by one device-scope release before its successors' counts.
    s << "struct TInf { long long tile0; int nb, DT; };\n"
         "__device__ __forceinline__ unsigned long long uni64_(unsigned long long v) {\n"
         "  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);\n"
         "  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));\n"
         "  return ((unsigned long long)hi << 32) | lo;\n}\n"
         "extern \"C\" __global__ void adp_tile_reset(const int* __restrict__ seed, int count, int ntiles,\n"
         "                                            unsigned long long* queue, int* ctrl, int* cnt) {\n"
         "  const int seeds = seed[count];\n"
         "  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < ntiles; s += gridDim.x * blockDim.x) {\n"
         "    cnt[s] = 0;\n"
         "    unsigned long long it = ~0ull;\n"
         "    if (s < seeds) {  // the diagonal tiles, instance by instance\n"
         "      int lo = 0, hi = count;\n"
         "      while (hi - lo > 1) { const int mid = (lo + hi) >> 1; if (seed[mid] <= s) lo = mid; else hi = mid; }\n"
         "      const unsigned long long d = (unsigned long long)(s - seed[lo]);\n"
         "      it = ((unsigned long long)lo << 40) | (d << 20) | d;\n"
         "    }\n"
         "    queue[s] = it;\n"
         "  }\n"
         "  if (blockIdx.x == 0 && threadIdx.x == 0) { ctrl[0] = 0; ctrl[1] = seeds; ctrl[2] = 0; }\n"
         "}\n";
    s << "// WH_: the whole input fits the tile's LDS copy (every input read is one\n"
         "// LDS read, no range tests); else the tile's windows i0 - 32 .. i0 + 95 and\n"
         "// j0 - 64 .. j0 + 63, and memory beyond them\n"
         "template <int NT_, bool WH_> __device__ __forceinline__ void tile_(const Inst& q, int TI_, int TJ_, MidSm& sm, input_t* tw,\n"
         "                                                               const RnaHead* rhp) {\n"
         "  const Dims z = q.z;\n  const Tabs T = q.T;\n"
         "  const int i0 = TI_ << 5, j0 = TJ_ << 5;\n"
         "  if (WH_) {\n"
         "    for (int e = threadIdx.x; e <= z.n1; e += NT_) tw[e] = ((const GLOBAL_ input_t*)q.in1)[e];\n"
         "  } else if (threadIdx.x < 256) {\n"
         "    const int x = threadIdx.x < 128 ? i0 - 32 + threadIdx.x : j0 - 64 + (threadIdx.x - 128);\n"
         "    const input_t v = ((const GLOBAL_ input_t*)q.in1)[min(max(x, 0), z.n1)];\n"
         "    if (threadIdx.x < 128) tinA_[threadIdx.x] = v; else tinB_[threadIdx.x - 128] = v;\n"
         "  }\n"
         "  __syncthreads();\n"
         "  const Ctx _x{q.in1, q.in2, z.n1, z.n2, (const RnaPar*)q.rna, q.hl, i0 - 32 + (1 << 30), j0 - 64 + (1 << 30),\n"
         "               WH_ ? max(z.n1 + 1, 1) : 0, (const LDS_ input_t*)tw, (const LDS_ RnaHead*)rhp};\n"
      << "  const int D = " << (m.window > 0 ? "min(" + std::to_string(m.window) + ", z.n1)" : std::string("z.n1")) << ";\n"
      << "  const int KLO = i0 + TB + " << mid_lo << ", KHI = j0 - 1 - " << mid_hi << ";\n"
      << "  const bool _mid = " << (has_mid ? "TJ_ - TI_ >= 2 && KLO <= KHI" : "false") << ";\n"
      // no middle product: an empty range far above every split (n < 2^30),
      // so _mhi + 1 + lane cannot overflow
      << "  const int _mlo = _mid ? KLO : (1 << 30), _mhi = _mid ? KHI : (1 << 30) - 1;\n";
    if (has_mid)
      s << "  if (_mid) {\n"
           "    middle_<1024 / NT_>(_x, z, T, D, i0, j0, _mlo, _mhi, sm);\n"
           "    __syncthreads();  // the partials are in the table entries\n"
           "  }\n";
    else
      s << "  (void)sm;\n";
    if (reg_h >= 0) {
      // the region's halo (finished neighbour tiles) into LDS; the tile's own
      // cells enter it as they are computed
      s << "  const int _ri0 = i0, _rj0 = j0 - RHALO_;\n"
           "  for (int e = threadIdx.x; e < RH_ * RW_; e += NT_) {\n"
           "    const int ra = e / RW_, rb = e - ra * RW_;\n"
           "    if (ra < TB && rb >= RHALO_) continue;\n"
           "    const long long x = sidx_(z, _ri0 + ra, _rj0 + rb);\n";
      for (int t : m.order) {
        s << "    tv_" << t << "[e] = ((const GLOBAL_ answer_t*)T.c[" << t << "])[x];\n";
        if (tile_rules.count(t)) s << "    trr_" << t << "[e] = ((const GLOBAL_ short*)T.r[" << t << "])[x];\n";
      }
      s << "  }\n";
      if (!strip_l.empty() || !strip_r.empty()) {
        s << "  for (int e = threadIdx.x; e < 2048; e += NT_) {\n";
        if (!strip_l.empty()) {
          s << "    {\n      const long long x = sidx_(z, i0 + (e >> 6), i0 + (e & 63));\n";
          for (int t : strip_l) {
            s << "      tv_" << t << "[SLO_" << t << " + e] = ((const GLOBAL_ answer_t*)T.c[" << t << "])[x];\n";
            if (tile_rules.count(t)) s << "      trr_" << t << "[SLO_" << t << " + e] = ((const GLOBAL_ short*)T.r[" << t << "])[x];\n";
          }
          s << "    }\n";
        }
        if (!strip_r.empty()) {
          s << "    {\n      const long long x = sidx_(z, j0 - 32 + (e >> 5), j0 + (e & 31));\n";
          for (int t : strip_r) {
            s << "      tv_" << t << "[SRO_" << t << " + e] = ((const GLOBAL_ answer_t*)T.c[" << t << "])[x];\n";
            if (tile_rules.count(t)) s << "      trr_" << t << "[SRO_" << t << " + e] = ((const GLOBAL_ short*)T.r[" << t << "])[x];\n";
          }
          s << "    }\n";
        }
        s << "  }\n";
      }
      s << "  __syncthreads();\n";
    } else {
      s << "  const int _ri0 = i0, _rj0 = j0;\n";
    }
    // the in-tile wavefront: tile diagonal e holds cells (r, r + e); a cell
    // reads cells of smaller e only, so one barrier per diagonal
    // lanes per cell: 8 (4 waves for a tile diagonal) -- the cell code is
    // long, so one active wave per SIMD beats more lanes per cell and the
    // other waves of a 1024-thread tile only meet the barriers; grammars
    // with nested split loops (interior loops) spread them over NT_ / 32
    bool nested = false;
    for (int t : m.order)
      for (auto& o : alts_tile[t]) nested = nested || o.loops.size() >= 2;
    s << "  constexpr int G = " << (!m.any_order() ? "1" : nested ? "NT_ / TB" : "8") << ";\n"
         "  const int grp = threadIdx.x / G, lane = threadIdx.x & (G - 1);\n"
         "  for (int e = TI_ == TJ_ ? 0 : 1 - TB; e < TB; ++e) {\n"
         "    const int r = grp + (e < 0 ? -e : 0), c = r + e;\n"
         "    const int i = i0 + r, j = j0 + c;\n"
         "    if (r < TB && c < TB && j <= z.n1 && j - i <= D) {\n";
    s << "      CellV cv;\n";
    for (int t : m.order) s << "      cv.r" << t << " = -1;\n";
    for (int t : m.order)
      s << "      tabT_" << t << "<G, false>(_x, z, T, i, j, lane, cv, _mid, _mlo, _mhi, _ri0, _rj0);\n";
    s << "    }\n"
         "    __syncthreads();\n"
         "  }\n";
    if (reg_h >= 0) {
      // the tile's records to memory: each table's 1024 entries contiguous
      s << "  {\n"
           "    const long long base = idx_(z, i0, j0);\n"
           "    for (int e = threadIdx.x; e < 1024; e += NT_) {\n"
           "      const int ra = e >> 5, rb = e & 31;\n";
      for (int t : m.order) {
        s << "      ((GLOBAL_ answer_t*)T.c[" << t << "])[base + e] = tv_" << t << "[ra * RW_ + rb + RHALO_];\n"
          << "      ((GLOBAL_ short*)T.r[" << t << "])[base + e] = tcr_" << t << "[e];\n";
        for (int q = 0; q < m.tabs[t].cat; ++q)
          s << "      ((GLOBAL_ int*)T.p[" << t << "][" << q << "])[base + e] = tcp_" << t << "_" << q << "[e];\n";
      }
      s << "    }\n"
           "  }\n";
    }
    s << "}\n";
    s << "template <int NT_> __device__ __forceinline__ void tile_run_(const Inst* __restrict__ I, const TInf* __restrict__ TI,\n"
         "                                                          int ntiles, unsigned long long* queue, int* ctrl, int* cnt) {\n"
         "  __shared__ MidSm sm;\n"
         "  __shared__ unsigned long long s_item;\n"
         "  // the whole input in LDS when it fits: 16 KB for a 1024-thread tile\n"
         "  // (few tiles per CU; 8 KB where the tables fill the LDS), 4 KB for\n"
         "  // 256 threads (several tiles per CU)\n"
         "  constexpr int WCAP = (NT_ == 1024 ? " << wcap_big << " : 4096) / (int)sizeof(input_t);\n"
         "  __shared__ input_t tinW[WCAP];\n"
         << (m.rna ? "  // the energy tables' small head, once per workgroup (every instance shares the tables)\n"
                     "  __shared__ RnaHead rhead;\n"
                     "  for (int e = threadIdx.x; e < (int)(sizeof(RnaHead) / 4); e += NT_) ((int*)&rhead)[e] = ((const GLOBAL_ int*)I[0].rna)[e];\n"
                     "  __syncthreads();\n"
                     "  const RnaHead* rhp = &rhead;\n"
                   : "  const RnaHead* rhp = nullptr;\n")
         << "  for (;;) {\n"
         "    if (threadIdx.x < 64) {  // wave 0 claims, with wave-uniform control flow\n"
         "      int slot = 0;\n"
         "      if (threadIdx.x == 0) slot = __hip_atomic_fetch_add(&ctrl[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "      slot = __builtin_amdgcn_readfirstlane(slot);\n"
         "      unsigned long long it = ~0ull;\n"
         "      if (slot < ntiles) {\n"
         "        it = uni64_(__hip_atomic_load(&queue[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));\n"
         "        for (int spin = 0; it == ~0ull; ++spin) {\n"
         "          if (spin > (1 << 23)) {  // a scheduling fault ends the launch with a flag, not a hang\n"
         "            if (threadIdx.x == 0) __hip_atomic_store(&ctrl[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "            break;\n"
         "          }\n"
         "          __builtin_amdgcn_s_sleep(4);\n"
         "          it = uni64_(__hip_atomic_load(&queue[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));\n"
         "        }\n"
         "        // no copy of another tile's lines cached here, from this launch\n"
         "        // or an earlier one, may be read in place of what its producer wrote\n"
         "        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, \"agent\");\n"
         "      }\n"
         "      if (threadIdx.x == 0) s_item = it;\n"
         "    }\n"
         "    __syncthreads();\n"
         "    const unsigned long long it = s_item;\n"
         "    __syncthreads();\n"
         "    if (it == ~0ull) break;\n"
         "    const int q = (int)(it >> 40), TI_ = (int)((it >> 20) & 0xfffff), TJ_ = (int)(it & 0xfffff);\n"
         "    if (I[q].z.n1 + 1 <= WCAP) tile_<NT_, true>(I[q], TI_, TJ_, sm, tinW, rhp);\n"
         "    else tile_<NT_, false>(I[q], TI_, TJ_, sm, tinW, rhp);\n"
         "    __builtin_amdgcn_fence(__ATOMIC_RELEASE, \"workgroup\");\n"
         "    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores are done\n"
         "    __syncthreads();\n"
         "    if (threadIdx.x == 0) {\n"
         "      __builtin_amdgcn_fence(__ATOMIC_RELEASE, \"agent\");\n"
         "      const TInf f = TI[q];\n"
         "      for (int s2 = 0; s2 < 2; ++s2) {  // successors (I - 1, J) and (I, J + 1)\n"
         "        const int I2 = s2 == 0 ? TI_ - 1 : TI_, J2 = s2 == 0 ? TJ_ : TJ_ + 1;\n"
         "        if (I2 < 0 || J2 >= f.nb || J2 - I2 > f.DT) continue;\n"
         "        const long long g = f.tile0 + tile_ix_(f.nb, I2, J2);\n"
         "        const int old = __hip_atomic_fetch_add(&cnt[g], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "        if (old == 1) {\n"
         "          const int pos = __hip_atomic_fetch_add(&ctrl[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "          __hip_atomic_store(&queue[pos], ((unsigned long long)q << 40) | ((unsigned long long)I2 << 20) | (unsigned long long)J2,\n"
         "                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "        }\n"
         "      }\n"
         "    }\n"
         "    __syncthreads();\n"
         "  }\n"
         "}\n"
         "extern \"C\" __global__ __launch_bounds__(256) void adp_tile_256(const Inst* I, const TInf* TI, int ntiles,\n"
         "    unsigned long long* queue, int* ctrl, int* cnt) { tile_run_<256>(I, TI, ntiles, queue, ctrl, cnt); }\n"
         "extern \"C\" __global__ __launch_bounds__(1024) void adp_tile_1024(const Inst* I, const TInf* TI, int ntiles,\n"
         "    unsigned long long* queue, int* ctrl, int* cnt) { tile_run_<1024>(I, TI, ntiles, queue, ctrl, cnt); }\n";
  }
  if (m.reg_band) {
    // The band pipeline below with the band's own cells in registers: every
    // reference is (i, j - dj) or (i - 1, j - dj) with dj small.  Lane r
    // keeps the cells it computed in the last steps (h<t>_k: k + 1 steps
    // ago); (i, j - dj) is its own h_(dj-1), (i - 1, j - dj) is lane r - 1's
    // h_dj (one lane shift), or for lane 0 the band above's last row, read
    // per chunk of BAND_CH columns into lanes 0..BAND_CH-1 after the acquire
    // and taken by readlane.  So no step waits for the stores of the one
    // before; the stores (tables for the backtrace and the next band) are
    // waited for only by the release before each publication.  The inputs
    // of the next step are loaded before this step's stores.
    std::map<int, int> H;   // history length per tabulation
    std::set<int> above;    // tabulations read from the row above
    for (auto& k : rset) {
      H[k[0]] = std::max(H[k[0]], k[2] + k[1]);
      if (k[1] == 1) above.insert(k[0]);
    }
    s << "#define BAND_CH 16\n"
         "extern \"C\" __global__ __launch_bounds__(64) void adp_band(const Inst* __restrict__ I, const int* __restrict__ boff,\n"
         "                                                           int count, int* ctrl, int* prog) {\n"
         "  const int lane = threadIdx.x;\n"
         "  const int total = boff[count];\n"
         "  for (;;) {\n"
         "    int tk = 0;\n"
         "    if (lane == 0) tk = __hip_atomic_fetch_add(&ctrl[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "    tk = __builtin_amdgcn_readfirstlane(tk);\n"
         "    if (tk >= total) break;\n"
         "    int lo = 0, hi = count;  // boff[lo] <= tk < boff[hi]\n"
         "    while (hi - lo > 1) { const int mid = (lo + hi) >> 1; if (boff[mid] <= tk) lo = mid; else hi = mid; }\n"
         "    const int q = __builtin_amdgcn_readfirstlane(lo);\n"
         "    const int band = tk - boff[q];\n"
         "    const Inst& In = I[q];\n"
         "    const Dims z = In.z;\n    const Tabs T = In.T;\n"
         "    const Ctx _x{In.in1, In.in2, z.n1, z.n2, (const RnaPar*)In.rna, In.hl};\n"
         "    const int i = band * 64 + lane, W = z.n2 + 1, steps = W + 63;\n"
         "    const bool row_ok = i <= z.n1;\n"
         "    int known = band == 0 ? W : 0;  // columns of the band above seen published\n"
         "    const int up = band * 64 - 1;  // last row of the band above\n"
         "    const long long xb = (long long)band * (z.n2 + 64) * 64 + lane;\n";
    for (auto& h : H)
      for (int k = 0; k < h.second; ++k) s << "    answer_t h" << h.first << "_" << k << "{}; short g" << h.first << "_" << k << " = -1;\n";
    for (int t : above) s << "    answer_t bv" << t << "{}, pv" << t << "{}; short br" << t << " = -1, pr" << t << " = -1;\n";
    // inputs: index i + c once per band; index j + c is what the lane above
    // had one step before, so it moves down one lane per step and lane 0
    // takes column st + c from the chunk's window (s0 + c + l on lane l)
    struct InRef { int t; std::string name; int track; char v; int c; };
    std::vector<InRef> ins;
    for (int t : m.order)
      for (auto& d : pre_decl_r[t]) {
        InRef r{t, d[1], 0, 'i', 0};
        char v = 0;
        if (std::sscanf(d[3].c_str(), "%d %c %d", &r.track, &v, &r.c) != 3) { err = "internal: input reference"; return false; }
        r.v = v;
        ins.push_back(r);
      }
    auto in_at = [&](const InRef& r, const std::string& x) {
      const std::string k = std::to_string(m.tracks == 1 ? 1 : r.track);
      return "((const GLOBAL_ input_t*)In.in" + std::to_string(r.track) + ")[min(max(" + x + ", 0), z.n" + k + ")]";
    };
    for (auto& r : ins) {
      const std::string nm = std::to_string(r.t) + r.name;
      if (r.v == 'i') s << "    const input_t bi" << nm << " = " << in_at(r, "i + " + std::to_string(r.c)) << ";\n";
      else
        s << "    input_t w1" << nm << "{}, x" << nm << " = " << in_at(r, "-1 - lane + " + std::to_string(r.c)) << ";  // step -1\n";
    }
    s << "    for (int s0 = 0; s0 < steps; s0 += BAND_CH) {\n"
         "      const int s1 = min(s0 + BAND_CH, steps), need = min(s1, W);\n"
         "      if (known < need) {\n"
         "        int k = __builtin_amdgcn_readfirstlane(__hip_atomic_load(&prog[tk - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));\n"
         "        for (int spin = 0; k < need; ++spin) {\n"
         "          if (spin > (1 << 24)) { if (lane == 0) __hip_atomic_store(&ctrl[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); k = need; break; }\n"
         "          __builtin_amdgcn_s_sleep(1);\n"
         "          k = __builtin_amdgcn_readfirstlane(__hip_atomic_load(&prog[tk - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));\n"
         "        }\n"
         "        known = k;\n"
         "        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, \"agent\");\n"
         "      }\n";
    if (!above.empty()) {
      s << "      if (band > 0) {  // columns s0 .. s0 + BAND_CH - 1 of the row above, one per lane\n";
      for (int t : above) s << "        pv" << t << " = bv" << t << "; pr" << t << " = br" << t << ";\n";
      s << "        if (lane < BAND_CH) {\n"
           "          const long long x = idx_(z, up, min(s0 + lane, W - 1));\n";
      for (int t : above)
        s << "          bv" << t << " = ((const GLOBAL_ answer_t*)T.c[" << t << "])[x]; br" << t << " = ((const GLOBAL_ short*)T.r[" << t << "])[x];\n";
      s << "        }\n"
           "      }\n";
    }
    for (auto& r : ins) {
      if (r.v != 'j') continue;
      const std::string nm = std::to_string(r.t) + r.name;
      s << "      w1" << nm << " = " << in_at(r, "s0 + " + std::to_string(r.c) + " + lane") << ";\n";
    }
    s << "      for (int st = s0; st < s1; ++st) {\n"
         "        const int j = st - lane;\n"
         "        Regs R;\n";
    for (auto& k : rset) {
      const std::string f = std::to_string(k[0]) + "_" + std::to_string(k[1]) + "_" + std::to_string(k[2]);
      const int t = k[0], di = k[1], dj = k[2];
      if (di == 0) {
        s << "        R.v" << f << " = h" << t << "_" << dj - 1 << "; R.r" << f << " = g" << t << "_" << dj - 1 << ";\n";
      } else {
        s << "        {\n"
             "          const int c = st - " << dj << " - s0, l = c >= 0 ? c : max(c + BAND_CH, 0);\n"
          << "          const answer_t bv = c >= 0 ? readlane_(bv" << t << ", l) : readlane_(pv" << t << ", l);\n"
          << "          const short br = c >= 0 ? readlane_(br" << t << ", l) : readlane_(pr" << t << ", l);\n"
          << "          const answer_t uv = shfl_up1_(h" << t << "_" << dj << ");\n"
          << "          const short ur = shfl_up1_(g" << t << "_" << dj << ");\n"
          << "          R.v" << f << " = lane == 0 ? bv : uv; R.r" << f << " = lane == 0 ? br : ur;\n"
          << "        }\n";
      }
    }
    s << "        CellV cv;\n";
    for (int t : m.order) s << "        cv.c" << t << " = answer_t{}; cv.r" << t << " = -1;\n";
    for (int t : m.order) s << "        PreR_" << t << " p" << t << ";\n        p" << t << ".pad_ = 0;\n";
    for (auto& r : ins) {
      const std::string nm = std::to_string(r.t) + r.name;
      if (r.v == 'i') {
        s << "        p" << r.t << "." << r.name << " = bi" << nm << ";\n";
      } else {
        s << "        {\n"
          << "          const input_t a = shfl_up1_(x" << nm << "), b = readlane_(w1" << nm << ", st - s0);\n"
          << "          x" << nm << " = lane == 0 ? b : a;\n"
          << "          p" << r.t << "." << r.name << " = x" << nm << ";\n"
          << "        }\n";
      }
    }
    s << "        if (row_ok && j >= 0 && j < W) {\n";
    for (int t : m.order) s << "          tabR_" << t << "<1, true>(_x, z, T, i, j, 0, cv, p" << t << ", R);\n";
    s << "          store_at(T, xb + (long long)st * 64, cv);  // idx_(z, i, j): r + j == st\n"
         "        }\n";
    for (auto& h : H) {
      for (int k = h.second - 1; k > 0; --k)
        s << "        h" << h.first << "_" << k << " = h" << h.first << "_" << k - 1 << "; g" << h.first << "_" << k << " = g" << h.first << "_" << k - 1 << ";\n";
      s << "        h" << h.first << "_0 = cv.c" << h.first << "; g" << h.first << "_0 = cv.r" << h.first << ";\n";
    }
    s << "      }\n"
         "      const int done = min(s1 - 63, W);  // lane 63 finished column s1 - 64\n"
         "      if (done > 0) {\n"
         "        __builtin_amdgcn_fence(__ATOMIC_RELEASE, \"agent\");\n"
         "        if (lane == 0) __hip_atomic_store(&prog[tk], done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "      }\n"
         "    }\n"
         "  }\n"
         "}\n";
  } else if (m.tracks == 2 && !m.loops) {
    // Two tracks read only cells (i' <= i, j' <= j), and without split loops
    // a cell is O(1) work: the whole table in one persistent launch instead of
    // one launch per anti-diagonal.  A wave owns a band of 64 rows (lane =
    // row) and sweeps the columns skewed by the lane (lane r is on column
    // s - r at step s), so every cell it reads from its own band is finished
    // by an earlier step; after each step the wave waits for its stores, which
    // the other lanes of the CU then read.  Rows of earlier bands come from
    // their publications: each band raises a progress word (columns whose 64
    // rows are done) every BAND_CH steps with an agent-scope release, and
    // acquires the band above's word before a chunk that reads past what it
    // has seen.  Bands are claimed in order by a ticket over (instance, band),
    // so a band only waits for a band that a running wave has claimed; the
    // waits are bounded and raise ctrl[1] instead of hanging.
    s << "#define BAND_CH 32\n"
         "extern \"C\" __global__ __launch_bounds__(64) void adp_band(const Inst* __restrict__ I, const int* __restrict__ boff,\n"
         "                                                           int count, int* ctrl, int* prog) {\n"
         "  const int lane = threadIdx.x;\n"
         "  const int total = boff[count];\n"
         "  for (;;) {\n"
         "    int tk = 0;\n"
         "    if (lane == 0) tk = __hip_atomic_fetch_add(&ctrl[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "    tk = __builtin_amdgcn_readfirstlane(tk);\n"
         "    if (tk >= total) break;\n"
         "    int lo = 0, hi = count;  // boff[lo] <= tk < boff[hi]\n"
         "    while (hi - lo > 1) { const int mid = (lo + hi) >> 1; if (boff[mid] <= tk) lo = mid; else hi = mid; }\n"
         "    const int q = __builtin_amdgcn_readfirstlane(lo);\n"
         "    const int band = tk - boff[q];\n"
         "    const Inst& In = I[q];\n"
         "    const Dims z = In.z;\n    const Tabs T = In.T;\n"
         "    const Ctx _x{In.in1, In.in2, z.n1, z.n2, (const RnaPar*)In.rna, In.hl};\n"
         "    const int i = band * 64 + lane, W = z.n2 + 1, steps = W + 63;\n"
         "    const bool row_ok = i <= z.n1;\n"
         "    int known = band == 0 ? W : 0;  // columns of the band above seen published\n"
         "    for (int s0 = 0; s0 < steps; s0 += BAND_CH) {\n"
         "      const int s1 = min(s0 + BAND_CH, steps), need = min(s1, W);\n"
         "      if (known < need) {\n"
         "        int k = __builtin_amdgcn_readfirstlane(__hip_atomic_load(&prog[tk - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));\n"
         "        for (int spin = 0; k < need; ++spin) {\n"
         "          if (spin > (1 << 24)) { if (lane == 0) __hip_atomic_store(&ctrl[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); k = need; break; }\n"
         "          __builtin_amdgcn_s_sleep(1);\n"
         "          k = __builtin_amdgcn_readfirstlane(__hip_atomic_load(&prog[tk - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));\n"
         "        }\n"
         "        known = k;\n"
         "        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, \"agent\");\n"
         "      }\n"
         "      for (int st = s0; st < s1; ++st) {\n"
         "        const int j = st - lane;\n"
         "        if (row_ok && j >= 0 && j < W) {\n";
    cell_body("          ", "1", "0", m.fixed_refs ? "true" : "false");
    if (m.fixed_refs) s << "          store_cell(z, T, i, j, cv);  // no reference of the cell reads what it writes\n";
    s << "        }\n"
         "        __builtin_amdgcn_s_waitcnt(0x0F70);\n"
         "        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, \"workgroup\");\n"
         "      }\n"
         "      const int done = min(s1 - 63, W);  // lane 63 finished column s1 - 64\n"
         "      if (done > 0) {\n"
         "        __builtin_amdgcn_fence(__ATOMIC_RELEASE, \"agent\");\n"
         "        if (lane == 0) __hip_atomic_store(&prog[tk], done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "      }\n"
         "    }\n"
         "  }\n"
         "}\n";
  }
  // finish: axiom cell (or best window), answer, backtrace; one wave per
  // instance, every lane running the same walk (lane 0 alone writes)
  const Tab& AX = m.tabs[m.axiom];
  s << "extern \"C\" __global__ __launch_bounds__(64) void adp_finish(const Inst* I) {\n"
       "  const int lane = threadIdx.x;\n"
       "  const Inst& q = I[blockIdx.x];\n"
       "  const Dims z = q.z;\n  const Tabs& T = q.T;\n"
       "  const Ctx _x{q.in1, q.in2, z.n1, z.n2, (const RnaPar*)q.rna, q.hl};\n"
       "  const input_t* _in1 = q.in1;\n  const input_t* _in2 = q.in2;\n  (void)_in1; (void)_in2;\n"
       "  answer_t* ans = q.ans;\n  int* out = q.out;\n  Rec* tr = q.tr;\n  const int cap = q.cap;\n"
       "  int i0, j0;\n  bool found;\n";
  if (m.window > 0) {
    s << "  found = false; i0 = 0; j0 = " << m.window << ";\n"
      << "  answer_t best{};\n"
      << "  for (int i = 0, j = " << m.window << "; j <= z.n1; ++i, ++j) {\n";
    if (!AX.always_valid) s << "    if (!" << Gen(m).valid(m.axiom, "i", "j") << ") continue;\n";
    s << "    const answer_t c = " << Gen(m).cost(m.axiom, "i", "j") << ";\n"
      << (m.h.k == A_SUM     ? "    if (found) best = answer_t(best + c); else { best = answer_t(answer_t(0) + c); i0 = i; j0 = j; found = true; }\n"
          : m.h.k == A_COUNT ? "    if (found) best = answer_t(best + 1); else { best = answer_t(1); i0 = i; j0 = j; found = true; }\n"
                             : "    if (!found || " + better("c", "best") + ") { best = c; i0 = i; j0 = j; found = true; }\n")
      << "  }\n"
      << "  if (lane == 0) *ans = best;  // zero when no window cell parses\n";
  } else {
    if (m.tracks == 1) s << "  i0 = 0; j0 = z.n1;\n";
    else s << "  i0 = z.n1; j0 = z.n2;\n";
    s << "  found = " << (AX.always_valid ? "true" : Gen(m).valid(m.axiom, "i0", "j0")) << ";\n"
      << "  if (lane == 0) *ans = " << Gen(m).cost(m.axiom, "i0", "j0") << ";\n";
  }
  s << "  if (lane == 0) { out[0] = found ? 1 : 0; out[2] = i0; out[3] = j0; }\n"
    << "  if (!tr || !found" << (m.additive() ? " || true) { if (lane == 0) out[1] = 0; return; }  // sums have no backtrace\n"
                                             : ") { if (lane == 0) out[1] = 0; return; }\n")
    << "  int wr = 0;\n"
       "  bool over = false;\n"
       // the queue's last 1024 records also in LDS: the walk reads them back
       // without a round trip through the trace in memory
       "  __shared__ int qi_[1024], qj_[1024], qt_[1024];\n"
       "  auto push = [&](int i, int j, int rule) {\n"
       "    if (wr < cap) {\n"
       "      if (lane == 0) {\n"
       "        tr[wr].i = i; tr[wr].j = j; tr[wr].rule = rule; tr[wr].pos[0] = 0; tr[wr].pos[1] = 0;\n"
       "        qi_[wr & 1023] = i; qj_[wr & 1023] = j; qt_[wr & 1023] = rule;\n"
       "      }\n"
       "      ++wr;\n"
       "    } else over = true;\n"
       "  };\n"
    << "  push(i0, j0, " << AX.id << ");\n"
    // the records queued so far are resolved a batch at a time: lane b
    // loads record rd + b's rule and splits (all in flight together), then
    // the walk takes them in queue order from the lanes; records queued
    // meanwhile wait for the next batch, so the order is the same
    << "  for (int rd = 0; rd < wr;) {\n"
       // lane 0's queue entries: the ring in LDS is ordered by the wave's
       // own LDS operations; entries read back from the trace in memory
       // wait for lane 0's stores
       "    if (wr - rd > 1024 - 64) __syncthreads();\n"
       "    const int nb = min(64, wr - rd);\n"
       "    int li = 0, lj = 0, lr = -1, lp0 = 0, lp1 = 0;\n"
       "    if (lane < nb) {\n"
       "      const int e = rd + lane;\n"
       "      const bool near = wr - e <= 1024;  // not yet overwritten in the ring\n"
       "      li = near ? qi_[e & 1023] : tr[e].i;\n"
       "      lj = near ? qj_[e & 1023] : tr[e].j;\n"
       "      const int tid = near ? qt_[e & 1023] : tr[e].rule;\n"
       "      const long long x = idx_(z, li, lj);\n"
       "      switch (tid) {\n";
  for (int t : m.order) {
    const Tab& T = m.tabs[t];
    s << "        case " << T.id << ": lr = T.r[" << t << "][x];";
    for (int q = 0; q < T.cat; ++q) s << " lp" << q << " = T.p[" << t << "][" << q << "][x];";
    s << " break;\n";
  }
  s << "      }\n"
       "    }\n"
       "    for (int b = 0; b < nb; ++b, ++rd) {\n"
       "      const int i = __builtin_amdgcn_readlane(li, b), j = __builtin_amdgcn_readlane(lj, b);\n"
       "      const int r = __builtin_amdgcn_readlane(lr, b);\n"
       "      const int p0 = __builtin_amdgcn_readlane(lp0, b), p1 = __builtin_amdgcn_readlane(lp1, b);\n"
       "      if (lane == 0) { tr[rd].rule = r; tr[rd].pos[0] = p0; tr[rd].pos[1] = p1; }\n"
       "      (void)i; (void)j; (void)p0; (void)p1;\n"
       "      switch (r) {\n";
  for (int t : m.order)
    for (auto& o : alts[t]) {
      s << "        case " << o.rule << ": {";
      for (size_t q = 0; q < o.var_of_pos.size(); ++q)
        if (!o.var_of_pos[q].empty()) s << " const int " << o.var_of_pos[q] << " = p" << q << ";";
      for (auto& kd : o.kids) s << " push(" << kd.a << ", " << kd.b << ", " << m.tabs[kd.tab].id << ");";
      s << " break; }\n";
    }
  s << "      }\n"
       "    }\n"
       "  }\n"
       "  if (lane == 0) out[1] = over ? -1 : wr;\n"
       "}\n";
  src = s.str();
  return true;
}

}  // namespace

// ------------------------------------------------------------------ compiled grammar

struct Grammar {
  Model m;
  std::string src;
  std::vector<char> code;
  hipModule_t mod = nullptr;
  hipFunction_t fill[3] = {nullptr, nullptr, nullptr}, finish = nullptr;  // fill: 1, 8, 64 lanes per cell
  hipFunction_t band = nullptr;  // two tracks without split loops: the band pipeline
  // one track: the tile queue (reset; persistent fill with 256 or 1024 threads per tile)
  hipFunction_t tile_reset = nullptr, tile[2] = {nullptr, nullptr};
  int cus = 0, occ[2] = {0, 0};  // compute units; resident tile workgroups per CU
  void* dev = nullptr;
  size_t dev_cap = 0;
  void* host = nullptr;  // pinned: answer, out[4], trace
  size_t host_cap = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};
};

Grammar* compile(const char* spec, bool load, std::string& err, int* code) {
  if (code) *code = DPX_EINVAL;
  if (!spec) { err = "null spec"; return nullptr; }
  std::unique_ptr<Grammar> g(new Grammar());
  J root;
  JParser jp{spec, spec, ""};
  if (!jp.val(root)) { err = "grammar JSON: " + jp.err; return nullptr; }
  Builder b(g->m);
  if (!b.build(root)) { err = "grammar: " + b.err; return nullptr; }
  if (!analyze(g->m, err)) { err = "grammar: " + err; return nullptr; }
  if (!generate(g->m, g->src, err)) { err = "grammar: " + err; return nullptr; }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, g->src.c_str(), "adp_grammar.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
    err = "hiprtcCreateProgram failed";
    return nullptr;
  }
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-fwrapv", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    err = "grammar does not compile (hipRTC):\n" + log;
    return nullptr;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  g->code.resize(cs);
  hiprtcGetCode(prog, g->code.data());
  hiprtcDestroyProgram(&prog);
  if (load) {
    if (code) *code = DPX_EHIP;
    if (hipModuleLoadData(&g->mod, g->code.data()) != hipSuccess ||
        (g->m.tracks == 2 && (hipModuleGetFunction(&g->fill[0], g->mod, "adp_fill_1") != hipSuccess ||
                              hipModuleGetFunction(&g->fill[1], g->mod, "adp_fill_8") != hipSuccess ||
                              hipModuleGetFunction(&g->fill[2], g->mod, "adp_fill_64") != hipSuccess)) ||
        hipModuleGetFunction(&g->finish, g->mod, "adp_finish") != hipSuccess ||
        (g->m.tracks == 2 && !g->m.loops && hipModuleGetFunction(&g->band, g->mod, "adp_band") != hipSuccess) ||
        (g->m.tracks == 1 && (hipModuleGetFunction(&g->tile_reset, g->mod, "adp_tile_reset") != hipSuccess ||
                              hipModuleGetFunction(&g->tile[0], g->mod, "adp_tile_256") != hipSuccess ||
                              hipModuleGetFunction(&g->tile[1], g->mod, "adp_tile_1024") != hipSuccess))) {
      err = "loading the compiled grammar failed";
      if (g->mod) (void)hipModuleUnload(g->mod);
      return nullptr;
    }
    if (g->m.tracks == 1) {
      int dev = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&g->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
          hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&g->occ[0], g->tile[0], 256, 0) != hipSuccess ||
          hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&g->occ[1], g->tile[1], 1024, 0) != hipSuccess) {
        err = "querying the tile kernels failed";
        (void)hipModuleUnload(g->mod);
        return nullptr;
      }
      g->cus = std::max(g->cus, 1);
      for (int& o : g->occ) o = std::max(o, 1);
    }
    if (hipEventCreate(&g->ev[0]) != hipSuccess || hipEventCreate(&g->ev[1]) != hipSuccess) {
      err = "event creation failed";
      (void)hipModuleUnload(g->mod);
      return nullptr;
    }
  }
  return g.release();
}

void destroy(Grammar* g) {
  if (!g) return;
  if (g->dev) (void)hipFree(g->dev);
  if (g->host) (void)hipHostFree(g->host);
  for (auto& e : g->ev)
    if (e) (void)hipEventDestroy(e);
  if (g->mod) (void)hipModuleUnload(g->mod);
  delete g;
}

const std::string& source(const Grammar* g) { return g->src; }

void info(const Grammar* g, dpx_grammar_info* o) {
  const Model& m = g->m;
  o->tracks = m.tracks;
  int used = 0;
  for (auto& t : m.tabs) used += t.used ? 1 : 0;
  o->n_tabs = used;
  o->n_rules = m.n_rules;
  o->max_pos = m.max_pos;
  o->axiom_id = m.tabs[m.axiom].id;
  o->window = m.window;
  o->answer_bytes = m.answer_bytes;
  o->alphabet_bytes = m.alphabet_bytes;
  o->loops = m.loops ? 1 : 0;
  o->rna = m.rna ? 1 : 0;
  o->code_bytes = (int64_t)g->code.size();
}

namespace {
size_t r256(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace

int parse_batch(Grammar* g, hipStream_t s, int count, const char* in1, const int64_t* off1, const char* in2,
                const int64_t* off2, void* answers, int32_t* found, dpx_trace_t* tr, const int64_t* tr_off,
                int32_t* tr_len, const void* rna, const int* hl, float* fill_ms, std::string& err) {
  const Model& m = g->m;
  if (!g->mod) { err = "grammar was compiled without loading"; return DPX_EINVAL; }
  if (m.rna && (!rna || !hl)) { err = "grammar calls librna functions: load an energy parameter file first"; return DPX_ENOPARAMS; }
  if (count == 0) return DPX_OK;
  if (count < 0 || !off1 || (m.tracks == 2 && !off2) || (tr && !tr_off)) { err = "bad batch arguments"; return DPX_EINVAL; }
  const int NT = (int)m.tabs.size();
  const size_t ab = (size_t)m.alphabet_bytes, ans4 = ((size_t)m.answer_bytes + 3) & ~(size_t)3;
  const size_t inst_bytes = 72 + 32 * (size_t)NT;
  struct Q { int64_t n1, n2, D, cells, nb = 0, DT = 0, tiles = 0; size_t o_in1, o_in2, o_tab, o_ans, o_tr; int64_t cap; };
  std::vector<Q> qs(count);
  size_t off = r256(inst_bytes * count);
  int64_t n1max = 0, cmax = 0, Dmax = 0;
  for (int q = 0; q < count; ++q) {
    Q& x = qs[q];
    x.n1 = off1[q + 1] - off1[q];
    x.n2 = m.tracks == 2 ? off2[q + 1] - off2[q] : 0;
    if (x.n1 < 0 || x.n2 < 0 || x.n1 > (1 << 30) || x.n2 > (1 << 30)) { err = "bad input offsets"; return DPX_EINVAL; }
    if (x.n1 > 0 && !in1) { err = "null input"; return DPX_EINVAL; }
    if (x.n2 > 0 && !in2) { err = "null input"; return DPX_EINVAL; }
    const int64_t W1 = x.n1 + 1;
    x.D = m.tracks == 1 ? (m.window > 0 ? std::min<int64_t>(m.window, x.n1) : x.n1) : x.n1 + x.n2;
    (void)W1;
    if (m.tracks == 1) {  // 32 x 32 tiles over the triangle (and the window)
      x.nb = (x.n1 + 32) / 32;
      x.DT = std::min<int64_t>(x.nb - 1, (x.D + 31) / 32);
      x.tiles = (x.DT + 1) * x.nb - x.DT * (x.DT + 1) / 2;
      x.cells = x.tiles * 1024;
    } else {
      x.cells = (x.n1 + 64) / 64 * 64 * (x.n2 + 64);
    }
    if (x.cells > ((int64_t)1 << 34)) { err = "grammar tables too large"; return DPX_EINVAL; }
    // one element of padding past each input: hoisted reads clamp their index to n
    x.o_in1 = off; off += r256((size_t)(x.n1 + 1) * ab + 16);
    x.o_in2 = off; off += r256((size_t)(x.n2 + 1) * ab + 16);
    x.o_tab = off;
    for (int t = 0; t < NT; ++t) {
      if (!m.tabs[t].used) continue;
      off += r256((size_t)x.cells * m.answer_bytes) + r256((size_t)x.cells * 2) + r256((size_t)x.cells * 4 * m.tabs[t].cat);
    }
    x.o_ans = off; off += r256(ans4 + 16);
    x.cap = tr ? tr_off[q + 1] - tr_off[q] : 0;
    if (x.cap < 0) { err = "bad trace offsets"; return DPX_EINVAL; }
    x.o_tr = off; off += r256(sizeof(dpx_trace_t) * (size_t)std::max<int64_t>(x.cap, 1));
    n1max = std::max(n1max, x.n1);
    cmax = std::max(cmax, m.tracks == 1 ? x.n1 + 1 : std::min(x.n1, x.n2) + 1);
    Dmax = std::max(Dmax, x.D);
  }
  // instance records go to the device largest D first
  std::vector<int> order(count);
  for (int q = 0; q < count; ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return qs[a].D > qs[b].D; });
  // band pipeline (two tracks without split loops): per record the first
  // band's ticket (count + 1 words), then the ticket and stall words, then
  // one progress word per band -- staged on the host with the zeros
  std::vector<int> band_blk;
  size_t o_band = 0, o_ctrl = 0;
  int64_t bands = 0;
  if (g->band) {
    for (int k = 0; k < count; ++k) bands += (qs[order[k]].n1 + 64) / 64;
    if (bands < ((int64_t)1 << 28)) {
      o_ctrl = r256((size_t)(count + 1) * 4) / 4;
      band_blk.assign(o_ctrl + 64 + (size_t)bands, 0);
      int64_t b = 0;
      for (int k = 0; k < count; ++k) {
        band_blk[k] = (int)b;
        b += (qs[order[k]].n1 + 64) / 64;
      }
      band_blk[count] = (int)b;
      o_band = off;
      off += r256(band_blk.size() * 4);
    }
  }
  // tile queue (one track): seeds and per-instance tile records (staged on
  // the host), then queue, dependency counts and control words (set by the
  // reset kernel, so the call holds no memset and can sit in a captured graph)
  std::vector<char> tile_blk;
  size_t o_tsched = 0, o_tinf = 0, o_queue = 0, o_cnt = 0, o_tctrl = 0;
  int64_t ntiles = 0;
  if (m.tracks == 1) {
    for (auto& x : qs) ntiles += x.tiles;
    if (ntiles >= ((int64_t)1 << 30) || count >= (1 << 23)) { err = "too many tiles in one call"; return DPX_EINVAL; }
    const size_t seed_bytes = r256((size_t)(count + 1) * 4);
    tile_blk.assign(seed_bytes + (size_t)count * 16, 0);
    int32_t* seeds = (int32_t*)tile_blk.data();
    int64_t tb = 0, sb = 0;
    for (int k = 0; k < count; ++k) {
      const Q& x = qs[order[k]];
      seeds[k] = (int32_t)sb;
      sb += x.nb;
      char* rec = tile_blk.data() + seed_bytes + (size_t)k * 16;
      const int32_t nd[2] = {(int32_t)x.nb, (int32_t)x.DT};
      std::memcpy(rec, &tb, 8);
      std::memcpy(rec + 8, nd, 8);
      tb += x.tiles;
    }
    seeds[count] = (int32_t)sb;
    o_tsched = off;
    o_tinf = off + seed_bytes;
    off += r256(tile_blk.size());
    o_queue = off; off += r256((size_t)ntiles * 8);
    o_cnt = off; off += r256((size_t)ntiles * 4);
    o_tctrl = off; off += 256;
  }
  if (off > g->dev_cap) {
    if (g->dev) (void)hipFree(g->dev);
    g->dev = nullptr;
    g->dev_cap = 0;
    if (hipMalloc(&g->dev, off) != hipSuccess) { err = "device allocation for the grammar tables failed"; return DPX_ENOMEM; }
    g->dev_cap = off;
  }
  // host staging: instance records, inputs, then results
  int64_t tr_total = 0;
  for (auto& x : qs) tr_total += x.cap;
  const size_t in_bytes = (size_t)(off1[count] - off1[0]) * ab + (m.tracks == 2 ? (size_t)(off2[count] - off2[0]) * ab : 0);
  const size_t res_bytes = (size_t)count * (ans4 + 16);
  const size_t hbytes = r256(inst_bytes * count) + r256(in_bytes + 16) + r256(res_bytes) +
                        r256(sizeof(dpx_trace_t) * (size_t)std::max<int64_t>(tr_total, 1)) + 256;
  if (hbytes > g->host_cap) {
    if (g->host) (void)hipHostFree(g->host);
    g->host = nullptr;
    g->host_cap = 0;
    if (hipHostMalloc(&g->host, hbytes, hipHostMallocDefault) != hipSuccess) { err = "pinned staging failed"; return DPX_ENOMEM; }
    g->host_cap = hbytes;
  }
  char* dev = (char*)g->dev;
  char* h = (char*)g->host;
  char* h_inst = h;
  char* h_res = h + r256(inst_bytes * count) + r256(in_bytes + 16);
  char* h_tr = h_res + r256(res_bytes);
  int* h_stall = (int*)(h_tr + r256(sizeof(dpx_trace_t) * (size_t)std::max<int64_t>(tr_total, 1)));
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess) err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipSuccess;
  };
  for (int k = 0; k < count; ++k) {
    const int q = order[k];
    const Q& x = qs[q];
    char* rec = h_inst + inst_bytes * k;
    std::memset(rec, 0, inst_bytes);
    auto put = [&](size_t at, const void* v, size_t n) { std::memcpy(rec + at, v, n); };
    void* p;
    p = dev + x.o_in1; put(0, &p, 8);
    p = dev + x.o_in2; put(8, &p, 8);
    p = dev + x.o_ans; put(16, &p, 8);
    p = dev + x.o_ans + ans4; put(24, &p, 8);
    p = x.cap > 0 ? dev + x.o_tr : nullptr; put(32, &p, 8);
    put(40, &rna, 8);
    put(48, &hl, 8);
    const int dims[4] = {(int)x.n1, (int)x.n2, (int)x.cap, 0};
    put(56, dims, 16);
    size_t o = x.o_tab;
    for (int t = 0; t < NT; ++t) {
      if (!m.tabs[t].used) continue;
      void* c = dev + o; o += r256((size_t)x.cells * m.answer_bytes);
      void* r = dev + o; o += r256((size_t)x.cells * 2);
      void* p0 = m.tabs[t].cat > 0 ? dev + o : nullptr;
      void* p1 = m.tabs[t].cat > 1 ? dev + o + (size_t)x.cells * 4 : nullptr;
      o += r256((size_t)x.cells * 4 * m.tabs[t].cat);
      put(72 + 8 * (size_t)t, &c, 8);
      put(72 + 8 * ((size_t)NT + t), &r, 8);
      put(72 + 8 * (2 * (size_t)NT + 2 * t), &p0, 8);
      put(72 + 8 * (2 * (size_t)NT + 2 * t + 1), &p1, 8);
    }
    if (x.n1 > 0 && !chk(hipMemcpyAsync(dev + x.o_in1, in1 + off1[q] * ab, (size_t)x.n1 * ab, hipMemcpyHostToDevice, s), "input copy")) return DPX_EHIP;
    if (x.n2 > 0 && !chk(hipMemcpyAsync(dev + x.o_in2, in2 + off2[q] * ab, (size_t)x.n2 * ab, hipMemcpyHostToDevice, s), "input copy")) return DPX_EHIP;
  }
  if (!chk(hipMemcpyAsync(dev, h_inst, inst_bytes * count, hipMemcpyHostToDevice, s), "instance copy")) return DPX_EHIP;
  const void* d_inst = dev;
  if (!chk(hipEventRecord(g->ev[0], s), "event")) return DPX_EHIP;
  // Diagonals: without split loops a diagonal of many cells gets its own
  // launch (one lane per cell) and runs of diagonals of at most 256 cells go
  // to one workgroup per instance that steps through them with a barrier in
  // between; with split loops every diagonal gets its own launch, with 1, 8
  // or 64 lanes per cell as the loops grow (about d split points on diagonal
  // d).  blockIdx.y picks the instance: those still active on diagonal d are
  // the first ones of the sorted records.
  // cells of diagonal d in the largest instance (a bound for every instance)
  auto cnt_bound = [&](int64_t d) -> int64_t {
    if (m.tracks == 1) return n1max + 1 - d;
    return std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(cmax, d + 1), Dmax - d + 1));
  };
  if (!band_blk.empty()) {
    // band offsets and zeroed counters by copy (no memset node, so the call
    // can sit in a captured graph)
    if (!chk(hipMemcpyAsync(dev + o_band, band_blk.data(), band_blk.size() * 4, hipMemcpyHostToDevice, s), "band setup")) return DPX_EHIP;
    const int* d_boff = (const int*)(dev + o_band);
    int* d_ctrl = (int*)(dev + o_band) + o_ctrl;
    int* d_prog = d_ctrl + 64;
    int cnt = count;
    void* bargs[] = {(void*)&d_inst, (void*)&d_boff, (void*)&cnt, (void*)&d_ctrl, (void*)&d_prog};
    const unsigned grid = (unsigned)std::min<int64_t>(bands, 2048);
    if (!chk(hipModuleLaunchKernel(g->band, grid, 1, 1, 64, 1, 1, 0, s, bargs, nullptr), "band launch")) return DPX_EHIP;
    if (!chk(hipMemcpyAsync(h_stall, d_ctrl + 1, 4, hipMemcpyDeviceToHost, s), "band status")) return DPX_EHIP;
  }
  if (m.tracks == 1) {
    if (!chk(hipMemcpyAsync(dev + o_tsched, tile_blk.data(), tile_blk.size(), hipMemcpyHostToDevice, s), "tile setup")) return DPX_EHIP;
    const int* d_seed = (const int*)(dev + o_tsched);
    const void* d_tinf = dev + o_tinf;
    unsigned long long* d_queue = (unsigned long long*)(dev + o_queue);
    int* d_cnt = (int*)(dev + o_cnt);
    int* d_tctrl = (int*)(dev + o_tctrl);
    int cnt = count, nt = (int)ntiles;
    void* rargs[] = {(void*)&d_seed, (void*)&cnt, (void*)&nt, (void*)&d_queue, (void*)&d_tctrl, (void*)&d_cnt};
    const unsigned rgrid = (unsigned)std::min<int64_t>((ntiles + 255) / 256, 4096);
    if (!chk(hipModuleLaunchKernel(g->tile_reset, rgrid, 1, 1, 256, 1, 1, 0, s, rargs, nullptr), "tile reset")) return DPX_EHIP;
    // 1024 threads per tile when few tiles can be ready at once (each tile
    // then runs fastest) or when the tile's LDS leaves room for one
    // workgroup per CU either way; 256 when the launch has many more ready
    // tiles than compute units
    int64_t ready = 0;
    for (auto& x : qs) ready += x.nb;
    int big = ready <= 2 * (int64_t)g->cus || g->occ[0] <= 1 ? 1 : 0;
    if (const char* e = std::getenv("DPX_ADP_TILE_THREADS")) big = std::atoi(e) == 1024 ? 1 : std::atoi(e) == 256 ? 0 : big;
    const unsigned grid = (unsigned)std::min<int64_t>(ntiles, (int64_t)g->cus * g->occ[big]);
    void* targs[] = {(void*)&d_inst, (void*)&d_tinf, (void*)&nt, (void*)&d_queue, (void*)&d_tctrl, (void*)&d_cnt};
    if (!chk(hipModuleLaunchKernel(g->tile[big], grid, 1, 1, big ? 1024 : 256, 1, 1, 0, s, targs, nullptr), "tile fill")) return DPX_EHIP;
    if (!chk(hipMemcpyAsync(h_stall, d_tctrl + 2, 4, hipMemcpyDeviceToHost, s), "tile status")) return DPX_EHIP;
  }
  int active = band_blk.empty() && m.tracks == 2 ? count : 0;
  for (int64_t d = 0; d <= Dmax && active > 0;) {
    while (active > 0 && qs[order[active - 1]].D < d) --active;
    const int64_t c = cnt_bound(d);
    int d1 = (int)d + 1;
    unsigned grid, block = 256;
    int gi = 0;  // lanes per cell: 1, 8, 64
    if (!m.loops && c <= 256) {
      while (d1 <= Dmax && cnt_bound(d1) <= 256) ++d1;
      grid = 1;
    } else {
      // lanes per cell: enough to spread long split loops over the chip
      // when few cells are on the diagonal (one large instance), one lane per
      // cell when a batch already fills it
      int64_t cells_d = 0;
      for (int k = 0; k < active; ++k) {
        const Q& x = qs[order[k]];
        cells_d += m.tracks == 1 ? x.n1 + 1 - d
                                 : std::max<int64_t>(0, std::min<int64_t>(d, x.n1) - std::max<int64_t>(0, d - x.n2) + 1);
      }
      gi = !m.loops || d < 24 || !m.any_order() ? 0 : (d >= 192 && cells_d * 64 <= ((int64_t)1 << 18)) ? 2 : (cells_d * 8 <= ((int64_t)1 << 19)) ? 1 : 0;
      const int64_t lanes = c * (gi == 0 ? 1 : gi == 1 ? 8 : 64);
      block = gi == 0 || lanes < 256 ? 64 : 256;
      grid = (unsigned)((lanes + block - 1) / block);
    }
    int d0 = (int)d;
    void* args[] = {(void*)&d_inst, (void*)&d0, (void*)&d1};
    if (!chk(hipModuleLaunchKernel(g->fill[gi], grid, (unsigned)active, 1, block, 1, 1, 0, s, args, nullptr), "fill launch")) return DPX_EHIP;
    d = d1;
  }
  if (!chk(hipEventRecord(g->ev[1], s), "event")) return DPX_EHIP;
  void* fargs[] = {(void*)&d_inst};
  if (!chk(hipModuleLaunchKernel(g->finish, (unsigned)count, 1, 1, 64, 1, 1, 0, s, fargs, nullptr), "finish launch")) return DPX_EHIP;
  for (int q = 0; q < count; ++q)
    if (!chk(hipMemcpyAsync(h_res + (ans4 + 16) * q, dev + qs[q].o_ans, ans4 + 16, hipMemcpyDeviceToHost, s), "result copy")) return DPX_EHIP;
  if (!chk(hipStreamSynchronize(s), "grammar run")) return DPX_EHIP;
  if (!band_blk.empty() && *h_stall) { err = "band pipeline stalled (a band waited too long for the band above)"; return DPX_EHIP; }
  if (m.tracks == 1 && *h_stall) { err = "tile queue stalled (a workgroup waited too long for a tile)"; return DPX_EHIP; }
  int rc = DPX_OK;
  int64_t tro = 0;
  for (int q = 0; q < count; ++q) {
    const int* hout = (const int*)(h_res + (ans4 + 16) * q + ans4);
    const int n_tr = hout[1];
    if (tr && qs[q].cap > 0 && n_tr > 0 &&
        !chk(hipMemcpyAsync(h_tr + sizeof(dpx_trace_t) * tro, dev + qs[q].o_tr, sizeof(dpx_trace_t) * (size_t)n_tr, hipMemcpyDeviceToHost, s), "trace copy"))
      return DPX_EHIP;
    tro += std::max(n_tr, 0);
  }
  if (tr && !chk(hipStreamSynchronize(s), "trace copy")) return DPX_EHIP;
  tro = 0;
  for (int q = 0; q < count; ++q) {
    const char* hr = h_res + (ans4 + 16) * q;
    const int* hout = (const int*)(hr + ans4);
    if (answers) std::memcpy((char*)answers + (size_t)m.answer_bytes * q, hr, m.answer_bytes);
    if (found) found[q] = hout[0];
    const int n_tr = hout[1];
    if (tr) {
      if (n_tr > 0) std::memcpy(tr + tr_off[q], h_tr + sizeof(dpx_trace_t) * tro, sizeof(dpx_trace_t) * (size_t)n_tr);
      if (tr_len) tr_len[q] = qs[q].cap > 0 ? n_tr : 0;
      if (qs[q].cap > 0 && n_tr < 0) { err = "trace capacity too small for instance " + std::to_string(q); rc = DPX_ETRACE; }
      tro += std::max(n_tr, 0);
    }
  }
  if (fill_ms) (void)hipEventElapsedTime(fill_ms, g->ev[0], g->ev[1]);
  return rc;
}

int parse(Grammar* g, hipStream_t s, const void* in1, int n1, const void* in2, int n2, void* answer, int* found,
          dpx_trace_t* tr, int cap, int* len, const void* rna, const int* hl, float* fill_ms, std::string& err) {
  if (n1 < 0 || n2 < 0) { err = "bad input"; return DPX_EINVAL; }
  const int64_t o1[2] = {0, n1}, o2[2] = {0, g->m.tracks == 2 ? n2 : 0}, to[2] = {0, tr && cap > 0 ? cap : 0};
  int32_t f = 0, l = 0;
  const int rc = parse_batch(g, s, 1, (const char*)in1, o1, (const char*)in2, g->m.tracks == 2 ? o2 : nullptr, answer, &f,
                             tr && cap > 0 ? tr : nullptr, tr && cap > 0 ? to : nullptr, &l, rna, hl, fill_ms, err);
  if (found) *found = f;
  if (len) *len = l;
  return rc;
}

}  // namespace adp
}  // namespace dpx
