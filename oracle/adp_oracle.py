"""CPU oracle for generic ADP grammars (test infrastructure only).

Used by tests/ as the parity reference of dpx_grammar_parse; nothing in the product
imports it.  It evaluates the same JSON grammar tree that the library
compiles, with the Python callables of the algebra functions, in plain
Python loops (small inputs only):

* analysis (BaseParsers.scala:95-142): tabulations reachable from the axiom;
  minimum yields by |rules| rounds of iteration, maximum yields by unfolding
  tabulations |rules| deep (`rmax`); evaluation order within a cell from the
  zero-yield dependencies; rule ids by tabulation name, `inner.alt` each;
* tabulation (CodeGen.scala:31-186): alternatives in Or order (Or lifted
  above Concat left-major, above Map and Filter), split points ascending
  with the outer concatenation outermost, a candidate kept only when
  strictly better by the signature's h (or when it is the cell's first);
  back-pointer (rule, split positions left to right, 0 for the unused ones);
* concatenation ranges (BaseParsers.scala:282-345): ~ with hints (1 on each
  side for p ~ p), -~ splitting track 1, ~- splitting track 2;
* backtrace (CodeGen.scala:213-252): BFS over (i, j, tabulation id) records;
* window (CodeGen.scala:325-348, ADPParsers.scala:9-13);
* nested aggregations (an aggregate below a concatenation or a map,
  CodeGen.scala:31-36,108-135): distributed over their Or alternatives like
  the reference's norm; each yields the first best candidate of its
  alternative (the Scala `aggr` over that list, BaseParsers.scala:228), whose
  split positions join the alternative's; the backtrace passes through them
  (genUnapp, CodeGen.scala:218);
* sum / count (BaseParsers.scala:145-153,385-392, the Scala `aggr`): the
  sum of the candidates' values (count: their number), a cell without a
  candidate has no answer; no backtrace (BaseParsers.scala:150: no candidate
  matches a sum); a window adds up the axiom's answers of every window
  (count: counts the windows that parse), ADPParsers.scala:11.

Independent checks of these semantics live beside it: `enumerate_all`
(every derivation's value, the reference's Scala `apply` without
aggregation, BaseParsers.scala:230-300) and `replay` (the reference's `build`,
BaseParsers.scala:178-183: a trace evaluated bottom-up through `reapply`).
"""
INF = -1

# integer answer types: (bits, signed) on gfx950 and x86-64
_INT_TYPES = {"int": (32, True), "int32_t": (32, True), "unsigned": (32, False), "unsigned int": (32, False),
              "uint32_t": (32, False), "long": (64, True), "long int": (64, True), "long long": (64, True),
              "int64_t": (64, True), "unsigned long": (64, False), "unsigned long long": (64, False),
              "uint64_t": (64, False), "size_t": (64, False), "short": (16, True), "int16_t": (16, True),
              "unsigned short": (16, False), "uint16_t": (16, False), "char": (8, True), "signed char": (8, True),
              "int8_t": (8, True), "unsigned char": (8, False), "uint8_t": (8, False)}


def _hm(a, b):
    return b if a == INF else a if b == INF else min(a, b)


class Oracle:
    def __init__(self, spec, funcs):
        """spec: the grammar dict (dpx.adp.Grammar.spec()); funcs: name -> callable."""
        self.s = spec
        self.f = funcs
        self.tracks = spec.get("tracks", 1)
        self.tabs = spec["tabulations"]
        self.axiom = spec["axiom"]
        self.h = spec["h"]
        self.window = spec.get("window", 0)
        self.empty = spec.get("empty")
        self.int_type = _INT_TYPES.get(spec.get("answer", "int"))
        self._analyze()

    def wrap(self, v):
        """A sum of integer answers wraps like the answer type (C with
        -fwrapv, the JVM's Int); other values are left alone."""
        if self.int_type is None or not isinstance(v, int):
            return v
        bits, signed = self.int_type
        v &= (1 << bits) - 1
        return v - (1 << bits) if signed and v >> (bits - 1) else v

    # ---- yields ----------------------------------------------------------------
    def _hint(self, n):
        if "hint" in n:
            return [INF if x == -1 else x for x in n["hint"]]
        if n.get("track", 0) == 0 and self._same(n["cat"][0], n["cat"][1]):
            return [1, INF, 1, INF]
        return [0, INF, 0, INF]

    def _same(self, a, b):
        if "tab" in a or "tab" in b:
            return a.get("tab") == b.get("tab") and "tab" in a and "tab" in b
        if "term" in a or "term" in b:
            return a.get("term") == b.get("term") and a.get("term") != "seq" and "term" in a and "term" in b
        return a == b

    def _min(self, n):
        if "tab" in n:
            return self.minv[n["tab"]]
        if "term" in n:
            return self._term_yield(n)[0]
        if "or" in n:
            return min(self._min(x) for x in n["or"])
        if "cat" in n:
            h = self._hint(n)
            return max(self._min(n["cat"][0]), h[0]) + max(self._min(n["cat"][1]), h[2])
        return self._min(self._inner(n))

    def _max(self, n):
        if "tab" in n:
            return self.maxv[n["tab"]]
        if "term" in n:
            return self._term_yield(n)[1]
        if "or" in n:
            v = [self._max(x) for x in n["or"]]
            return INF if INF in v else max(v)
        if "cat" in n:
            h = self._hint(n)
            a, b = _hm(self._max(n["cat"][0]), h[1]), _hm(self._max(n["cat"][1]), h[3])
            return INF if INF in (a, b) else a + b
        return self._max(self._inner(n))

    def _rmax(self, n, d):
        if "tab" in n:
            return INF if d == 1 else self._rmax(self.tabs[n["tab"]]["p"], d - 1)
        if "term" in n:
            return self._term_yield(n)[1]
        if "or" in n:
            v = []
            for x in n["or"]:
                y = self._rmax(x, d)
                if y == INF:
                    return INF
                v.append(y)
            return max(v)
        if "cat" in n:
            a = self._rmax(n["cat"][0], d)
            if a == INF:
                return INF
            b = self._rmax(n["cat"][1], d)
            return INF if b == INF else a + b
        return self._rmax(self._inner(n), d)

    @staticmethod
    def _inner(n):
        for k in ("map", "filter", "agg"):
            if k in n:
                return n[k]
        raise ValueError("bad node %r" % n)

    def _term_yield(self, n):
        t = n["term"]
        if t in ("empty", "emptyi"):
            return 0, 0
        if t in ("empty1", "empty2"):
            return 0, INF
        if t == "seq":
            return n.get("min", 1), n.get("max", INF)
        return 1, 1

    def _alt(self, n):
        if "tab" in n or "term" in n:
            return 1
        if "or" in n:
            return sum(self._alt(x) for x in n["or"])
        if "cat" in n:
            return self._alt(n["cat"][0]) * self._alt(n["cat"][1])
        return self._alt(self._inner(n))

    def _has_bt(self, n):
        l, r = n["cat"]
        a, b, c, d = self._min(l), self._max(l), self._min(r), self._max(r)
        t = n.get("track", 0)
        if t == 0:
            return not ((a == b and a >= 0) or (c == d and c >= 0))
        if t == 1:
            return not (a == b and a >= 0)
        return not (c == d and c >= 0)

    def _cat(self, n):
        if "tab" in n or "term" in n:
            return 0
        if "or" in n:
            return max(self._cat(x) for x in n["or"])
        if "cat" in n:
            return self._cat(n["cat"][0]) + (1 if self._has_bt(n) else 0) + self._cat(n["cat"][1])
        return self._cat(self._inner(n))

    def _indices(self, n):
        h = self._hint(n)
        l, r = n["cat"]
        return max(self._min(l), h[0]), _hm(self._max(l), h[1]), max(self._min(r), h[2]), _hm(self._max(r), h[3])

    def _analyze(self):
        used = []

        def use(n):
            if "tab" in n:
                if n["tab"] not in used:
                    used.append(n["tab"])
                    use(self.tabs[n["tab"]]["p"])
                return
            for k in ("or", "cat"):
                if k in n:
                    for x in n[k]:
                        use(x)
                    return
            if "term" not in n:
                use(self._inner(n))

        use({"tab": self.axiom})
        self.used = [t for t in self.tabs if t in used]  # spec order
        R = len(self.used)
        self.minv = {t: 100000 for t in self.used}
        self.maxv = {}
        for _ in range(R):
            for t in self.used:
                self.minv[t] = self._min(self.tabs[t]["p"])
        for t in self.used:
            self.maxv[t] = self.minv[t]
        for t in self.used:
            self.maxv[t] = self._rmax(self.tabs[t]["p"], R)

        def deps(n, out):
            if "tab" in n:
                out.add(n["tab"])
            elif "or" in n:
                for x in n["or"]:
                    deps(x, out)
            elif "cat" in n:
                iL, _, jL, _ = self._indices(n)
                if iL == 0:
                    deps(n["cat"][1], out)
                if jL == 0:
                    deps(n["cat"][0], out)
            elif "term" not in n:
                deps(self._inner(n), out)

        dep = {}
        for t in self.used:
            dep[t] = set()
            deps(self.tabs[t]["p"], dep[t])
        self.order = []
        while len(self.order) < R:
            ready = [t for t in self.used if t not in self.order and dep[t] <= set(self.order)]
            if not ready:
                raise ValueError("loop between tabulations")
            self.order += ready
        self.id, self.alts = {}, {}
        k = 0
        for t in sorted(self.used):
            self.id[t] = k
            k += self._alt(self.tabs[t]["p"])
        self.n_rules = k
        self.catn = {t: self._cat(self.tabs[t]["p"]) for t in self.used}
        if max(self.catn.values()) > 2:  # trace records hold two split positions (as the library)
            raise ValueError("a rule needs more than 2 split positions")
        for t in self.used:
            self.alts[t] = self._flatten(self.tabs[t]["p"])
        self.tab_of_rule = {}
        for t in self.used:
            for q in range(len(self.alts[t])):
                self.tab_of_rule[self.id[t] + q] = (t, q)

    def _flatten(self, n, top=True):
        if "or" in n:
            return [a for x in n["or"] for a in self._flatten(x, top)]
        if "agg" in n:
            if top:  # the tabulation's own aggregation
                return self._flatten(n["agg"], True)
            # nested: one copy per inner alternative, numbered like the
            # reference's norm; each aggregates the union of them (the Scala
            # parsers' semantics) and counts only when its own alternative wins
            sibs = self._flatten(n["agg"], False)
            return [dict(n, agg=a, _sibs=sibs, _k=q) for q, a in enumerate(sibs)]
        if "map" in n:
            return [dict(n, map=a) for a in self._flatten(n["map"], False)]
        if "filter" in n:
            return [dict(n, filter=a) for a in self._flatten(n["filter"], top)]
        if "cat" in n:
            return [dict(n, cat=[a, b]) for a in self._flatten(n["cat"][0], False) for b in self._flatten(n["cat"][1], False)]
        return [n]

    # ---- aggregation order ------------------------------------------------------
    def _key(self, v):
        h = self.h
        if isinstance(h, dict):
            return self.f[list(h.values())[0]](v)
        return v

    def _additive(self):
        return self.h in ("sum", "count")

    def _better(self, c, c0, h=None):
        h = self.h if h is None else h
        kind = h if isinstance(h, str) else list(h.keys())[0]
        key = (lambda v: v) if isinstance(h, str) else self.f[list(h.values())[0]]
        if kind in ("max", "maxBy"):
            return key(c) > key(c0)
        return key(c) < key(c0)

    def _first_best(self, cands, h):
        """The first best of a list of candidates (value, positions, kids)
        by aggregation h, or None for an empty list."""
        best = None
        for c in cands:
            if best is None or self._better(c[0][0], best[0][0], h):
                best = c
        return best

    # ---- evaluation -------------------------------------------------------------
    def _cands(self, n, a, b, n1, n2):
        """(values, positions, kids) of every candidate of an Or-free node
        over subword (a, b), in loop order."""
        if "tab" in n:
            t = n["tab"]
            if self.tracks == 1:
                mn, mx = self.minv[t], self.maxv[t]
                if mn == mx:
                    if a + mn != b:
                        return
                else:
                    if mx != INF and b - mx > a:
                        return
                    if mn > 0 and a + mn > b:
                        return
            if not self.tabs[t].get("always_valid", False) and not self._valid(t, a, b):
                return
            yield [self.cost[t][(a, b)]], [], [(t, a, b)]
            return
        if "term" in n:
            t = n["term"]
            if t == "empty":
                if self.tracks == 2 and (a != 0 or b != 0):
                    return
                if a == b:
                    yield [], [], []
            elif t == "emptyi":
                if a == b:
                    yield [a], [], []
            elif t == "empty1":
                if a == 0:
                    yield [b], [], []
            elif t == "empty2":
                if b == 0:
                    yield [a], [], []
            elif t in ("el", "el1"):
                if a + 1 == b:
                    yield [self.in1[a]], [], []
            elif t == "el2":
                if a + 1 == b:
                    yield [self.in2[a]], [], []
            elif t == "eli":
                if a + 1 == b:
                    yield [a], [], []
            elif t == "seq":
                mn, mx = n.get("min", 1), n.get("max", INF)
                ok = (a + mn == b) if mn == mx else ((mx == INF or b - mx <= a) and (mn <= 0 or a + mn <= b))
                if ok:
                    yield [(a, b)], [], []
            return
        if "filter" in n:
            if not self.f[n["f"]](a, b):
                return
            yield from self._cands(n["filter"], a, b, n1, n2)
            return
        if "agg" in n:  # a nested aggregation: the first best of its alternatives' union
            if n["h"] in ("sum", "count"):
                raise ValueError("nested sum/count aggregations are not supported")
            sibs, k = n.get("_sibs", [n["agg"]]), n.get("_k", 0)
            best, won = None, -1
            for q, alt in enumerate(sibs):
                for c in self._cands(alt, a, b, n1, n2):
                    if len(c[0]) != 1:
                        raise ValueError("nested aggregation without one value")
                    if best is None or self._better(c[0][0], best[0][0], n["h"]):
                        best, won = c, q
            if best is not None and won == k:
                yield best
            return
        if "map" in n:
            for v, p, k in self._cands(n["map"], a, b, n1, n2):
                yield [self.f[n["f"]](*v)], p, k
            return
        if "cat" in n:
            l, r = n["cat"]
            iL, iU, jL, jU = self._indices(n)
            bt = self._has_bt(n)
            tr = n.get("track", 0)
            if tr == 0:
                lo, hi = a + iL, b - jL
                if jU != INF:
                    lo = max(lo, b - jU)
                if iU != INF:
                    hi = min(hi, a + iU)
                ks = range(lo, hi + 1)
                sub = lambda k: ((a, k), (k, b))
            else:
                lo_, up, v = (iL, iU, a) if tr == 1 else (jL, jU, b)
                ks = range(max(0, v - up) if up != INF else 0, v - lo_ + 1)
                sub = (lambda k: ((k, a), (k, b))) if tr == 1 else (lambda k: ((a, k), (k, b)))
            for k in ks:
                (la, lb), (ra, rb) = sub(k)
                for lv, lp, lk in self._cands(l, la, lb, n1, n2):
                    for rv, rp, rk in self._cands(r, ra, rb, n1, n2):
                        yield lv + rv, lp + ([k] if bt else []) + rp, lk + rk
            return
        raise ValueError("unexpected node %r" % n)

    def _valid(self, t, a, b):
        if self.empty is not None:
            return self.cost[t][(a, b)] != self.empty_value
        return self.rule[t][(a, b)] != -1

    def _cells(self, n1, n2):
        if self.tracks == 1:
            D = min(self.window, n1) if self.window > 0 else n1
            for d in range(D + 1):
                for i in range(n1 + 1 - d):
                    yield i, i + d
        else:
            for s in range(n1 + n2 + 1):
                for i in range(max(0, s - n2), min(n1, s) + 1):
                    yield i, s - i

    def run(self, in1, in2=None, trace=True, zero_value=0):
        """(answer, found, trace records (i, j, rule, pos0, pos1))."""
        self.in1, self.in2 = list(in1), list(in2 or [])
        n1, n2 = len(self.in1), len(self.in2)
        self.empty_value = self.empty if self.empty is None else _literal(self.empty)
        self.cost = {t: {} for t in self.used}
        self.rule = {t: {} for t in self.used}
        self.pos = {t: {} for t in self.used}
        for (i, j) in self._cells(n1, n2):
            for t in self.order:
                c0 = self.empty_value if self.empty is not None else zero_value
                r0, p0 = -1, [0, 0]
                for q, alt in enumerate(self.alts[t]):
                    for v, p, _ in self._cands(alt, i, j, n1, n2):
                        if len(v) != 1:
                            raise ValueError("alternative without one value")
                        c = v[0]
                        if self._additive():
                            inc = 1 if self.h == "count" else c
                            if r0 == -1:
                                c0, r0, p0 = self.wrap(inc), self.id[t] + q, (p + [0, 0])[:2]
                            else:
                                c0 = self.wrap(c0 + inc)
                        elif self._better(c, c0) or (self.empty is None and r0 == -1):
                            c0, r0, p0 = c, self.id[t] + q, (p + [0, 0])[:2]
                self.cost[t][(i, j)] = c0
                self.rule[t][(i, j)] = r0
                self.pos[t][(i, j)] = p0
        ax = self.axiom
        if self.window > 0:
            found, best, i0, j0 = False, None, 0, self.window
            i, j = 0, self.window
            while j <= n1:
                if self.tabs[ax].get("always_valid", False) or self._valid(ax, i, j):
                    c = self.cost[ax][(i, j)]
                    if self._additive():
                        inc = 1 if self.h == "count" else c
                        best = self.wrap(best + inc) if found else self.wrap(inc)
                        if not found:
                            i0, j0, found = i, j, True
                    elif not found or self._better(c, best):
                        best, i0, j0, found = c, i, j, True
                i, j = i + 1, j + 1
            ans = best
        else:
            i0, j0 = (0, n1) if self.tracks == 1 else (n1, n2)
            found = self.tabs[ax].get("always_valid", False) or self._valid(ax, i0, j0)
            ans = self.cost[ax][(i0, j0)]
        if not trace or not found or self._additive():
            return ans, found, [] if trace else None
        tr = [[i0, j0, self.id[ax], 0, 0]]
        rd = 0
        while rd < len(tr):
            i, j, tid = tr[rd][0], tr[rd][1], tr[rd][2]
            t = self.tab_of_rule[tid][0]
            r = self.rule[t][(i, j)]
            tr[rd][2] = r
            p = self.pos[t][(i, j)]
            tr[rd][3:5] = [p[q] if q < self.catn[t] else 0 for q in range(2)]
            if r >= 0:
                t2, q = self.tab_of_rule[r]
                for kid in self._kids(self.alts[t2][q], i, j, tr[rd][3:5]):
                    tr.append([kid[1], kid[2], self.id[kid[0]], 0, 0])
            rd += 1
        return ans, found, [tuple(x) for x in tr]

    def _split(self, n, a, b, pos, base):
        """Subwords of both sides of a concatenation in the backtrace: the
        split from the record's positions when the concatenation has one,
        else from the fixed side (genUnapp, CodeGen.scala:222-229)."""
        l, r = n["cat"]
        cl = self._cat(l)
        bt = self._has_bt(n)
        iL, iU, jL, jU = self._indices(n)
        tr = n.get("track", 0)
        if bt:
            k = pos[base + cl]
        elif tr == 0:
            k = a + iL if iL == iU else b - jL
        elif tr == 1:
            k = a - iL
        else:
            k = b - jL
        nb = base + cl + (1 if bt else 0)
        if tr == 1:
            return (k, a), (k, b), nb
        return (a, k), (k, b), nb

    def _kids(self, n, a, b, pos):
        """Tabulation leaves (tab, i, j) of an alternative with its split
        positions, left to right (genUnapp)."""
        out = []

        def walk(n, a, b, base):
            if "tab" in n:
                out.append((n["tab"], a, b))
            elif "cat" in n:
                (la, lb), (ra, rb), nb = self._split(n, a, b, pos, base)
                walk(n["cat"][0], la, lb, base)
                walk(n["cat"][1], ra, rb, nb)
            elif "term" not in n:
                walk(self._inner(n), a, b, base)

        walk(n, a, b, 0)
        return out

    # ---- independent checks -----------------------------------------------------
    def enumerate_all(self, in1, in2=None, limit=200000, aggregate_tabs=False):
        """Values of every derivation of the axiom (no aggregation): the
        reference's Scala `apply` semantics on the original tree.  With
        aggregate_tabs, a tabulation referenced inside a derivation yields
        its aggregated list (its first best by h), as a Scala Tabulate
        memoises (BaseParsers.scala:200-212): the semantics a nested
        aggregation with another h than the tabulations it reads sees."""
        self.in1, self.in2 = list(in1), list(in2 or [])
        n1, n2 = len(self.in1), len(self.in2)
        memo = {}
        count = [0]

        def ap(n, a, b, top=True):
            if "tab" in n:
                key = (n["tab"], a, b)
                if key not in memo:
                    memo[key] = None  # a cycle yields nothing
                    memo[key] = ap(self.tabs[n["tab"]]["p"], a, b)
                vs = memo[key] or []
                if aggregate_tabs and vs and not top:
                    best = None
                    for v in vs:
                        if best is None or self._better(_value(v), _value(best)):
                            best = v
                    vs = [best]
                return vs
            if "or" in n:
                return [v for x in n["or"] for v in ap(x, a, b, top)]
            if "agg" in n:
                vs = ap(n["agg"], a, b, top)
                if top or not vs:  # the tabulation's aggregation: every derivation
                    return vs
                if n["h"] in ("sum", "count"):
                    raise ValueError("nested sum/count aggregations are not supported")
                # nested (BaseParsers.scala:228): the first best of its list
                best = None
                for v in vs:
                    x = _value(v)
                    if best is None or self._better(x, _value(best), n["h"]):
                        best = v
                return [best]
            if "filter" in n:
                return ap(n["filter"], a, b, top) if self.f[n["f"]](a, b) else []
            if "map" in n:
                return [self.f[n["f"]](*_flat(v)) for v in ap(n["map"], a, b, False)]
            if "term" in n:
                return [v[0] if v else None for v, _, _ in self._cands(n, a, b, n1, n2)]
            l, r = n["cat"]
            iL, iU, jL, jU = self._indices(n)
            tr = n.get("track", 0)
            out = []
            if tr == 0:
                lo = a + iL if (jU == INF or a + iL > b - jU) else b - jU
                hi = b - jL if (iU == INF or b - jL < a + iU) else a + iU
                for k in range(lo, hi + 1):
                    L, Rr = ap(l, a, k, False), ap(r, k, b, False)
                    out += [_Pair(x, y) for x in L for y in Rr]
            elif tr == 1:
                lo = 0 if (iU == INF or a - iU < 0) else a - iU
                for k in range(lo, a - iL + 1):
                    L, Rr = ap(l, k, a, False), ap(r, k, b, False)
                    out += [_Pair(x, y) for x in L for y in Rr]
            else:
                lo = 0 if (jU == INF or b - jU < 0) else b - jU
                for k in range(lo, b - jL + 1):
                    L, Rr = ap(l, a, k, False), ap(r, k, b, False)
                    out += [_Pair(x, y) for x in L for y in Rr]
            count[0] += len(out)
            if count[0] > limit:
                raise OverflowError("too many derivations for brute force")
            return out

        if self.window > 0:
            vals = []
            for i in range(0, n1 - self.window + 1):
                vals += [v for v in ap({"tab": self.axiom}, i, i + self.window)]
            return vals
        return ap({"tab": self.axiom}, 0 if self.tracks == 1 else n1, n1 if self.tracks == 1 else n2)

    def replay(self, in1, in2, trace):
        """The answer a trace stands for (the reference's `build`): records
        evaluated from the last to the first, each by its rule and split
        positions, tabulation leaves taken from records evaluated before."""
        self.in1, self.in2 = list(in1), list(in2 or [])
        got = {}
        val = None
        for (i, j, r, p0, p1) in reversed(trace):
            t, q = self.tab_of_rule[r]
            alt = self.alts[t][q]
            val = self._eval_fixed(alt, i, j, [p0, p1], got)
            got[(t, i, j)] = val
        return val

    def _eval_fixed(self, n, a, b, pos, got):
        vals = []

        def ev(n, a, b, base):
            if "tab" in n:
                return [got[(n["tab"], a, b)]]
            if "term" in n:
                c = list(self._cands(n, a, b, 0, 0))
                if not c:
                    raise ValueError("terminal does not match in replay")
                return c[0][0]
            if "filter" in n:
                if not self.f[n["f"]](a, b):
                    raise ValueError("filter fails in replay")
                return ev(n["filter"], a, b, base)
            if "map" in n:
                return [self.f[n["f"]](*ev(n["map"], a, b, base))]
            if "agg" in n:  # the record's positions name the nested best
                return ev(n["agg"], a, b, base)
            (la, lb), (ra, rb), nb = self._split(n, a, b, pos, base)
            return ev(n["cat"][0], la, lb, base) + ev(n["cat"][1], ra, rb, nb)

        vals = ev(n, a, b, 0)
        return vals[0]


class _Pair:
    __slots__ = ("x", "y")

    def __init__(self, x, y):
        self.x, self.y = x, y


def _value(v):
    """The value of a derivation element (a mapped value; a lone leaf)."""
    f = _flat(v)
    return f[0] if len(f) == 1 else tuple(f)


def _flat(v):
    if isinstance(v, _Pair):
        return [a for a in _flat(v.x) + _flat(v.y) if a is not _UNIT]
    return [v] if v is not None else []


_UNIT = object()


def _literal(s):
    try:
        return int(s, 0)
    except ValueError:
        return float(s)
