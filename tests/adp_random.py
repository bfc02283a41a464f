"""Seeded random ADP grammars for the generic path's parity tests: every
construct the library accepts (terminals incl. bounded seq, concatenations
with and without hints, -~ and ~-, filters, maps, tabulation references,
max/min/maxBy, always_valid and empty values, windows), with C bodies and
Python callables of the same meaning."""
import random

import dpx.adp as A

HEADER = ("__device__ __forceinline__ int val_(int x) { return x; }\n"
          "__device__ __forceinline__ int val_(char x) { return x; }\n"
          "__device__ __forceinline__ int val_(T2ii s) { return 2 * s._1 - s._2; }\n")


def _val(x):
    return 2 * x[0] - x[1] if isinstance(x, tuple) else x


def _mapfn(rng, nargs):
    # non-negative weights: every algebra function is monotone in its answers,
    # so the DP optimum is the optimum over all derivations (Bellman)
    w = [rng.choice([0, 1, 1, 2, 3]) for _ in range(nargs)]
    c = rng.randint(-3, 5)
    body = "return %d%s;" % (c, "".join(" + (%d) * val_(a%d)" % (w[q], q) for q in range(nargs)))
    args = ",".join("a%d" % q for q in range(nargs))
    return A.cfun(lambda *a, w=w, c=c: _wrap(c + sum(w[q] * _val(a[q]) for q in range(nargs))), args, body)


def _mulfn(rng, nargs):
    # products of the arguments' values and a weight: with h = sum the DP
    # answer is the sum of every derivation's value (distributivity, exact in
    # int32 arithmetic modulo 2^32)
    w = rng.choice([1, 1, 2, 3, -1])
    body = "return %d%s;" % (w, "".join(" * val_(a%d)" % q for q in range(nargs)))
    args = ",".join("a%d" % q for q in range(nargs))

    def f(*a, w=w):
        v = w
        for x in a:
            v *= _val(x)
        return _wrap(v)
    return A.cfun(f, args, body)


def _wrap(v):
    return (v + (1 << 31)) % (1 << 32) - (1 << 31)


def single_track(seed, agg=None):
    """agg "sum" or "count": that aggregation with product algebras."""
    rng = random.Random(seed)
    h = rng.choice(["max", "min", "maxBy"])
    mapf = _mapfn
    key = None
    if agg:
        h, mapf = agg, _mulfn
    elif h == "maxBy":
        key = A.cfun(lambda a: a % 5 if a >= 0 else -((-a) % 5), "a", "return a % 5;")  # C remainder: many ties
        h = A.max_by(key)
    window = rng.choice([0, 0, 0, 4, 6])
    empty = None
    if isinstance(h, str) and not agg and rng.random() < 0.3:  # the reference's cudaEmpty: never accepted as a candidate
        empty = "-1000000" if h == "max" else "1000000"
    g = A.Grammar(tracks=1, h=h, window=window, header=HEADER, empty_value=empty)
    agg = h
    ntab = rng.randint(1, 3)
    names = ["T%d" % q for q in range(ntab)]
    tabs = {nm: g.tabulate(nm, None) for nm in names}  # parsers set below

    def item(depth):
        r = rng.random()
        if r < 0.35:
            return tabs[rng.choice(names)]
        if r < 0.5:
            return g.el
        if r < 0.6:
            return g.eli
        if r < 0.7:
            return g.emptyi
        if r < 0.8:
            mn = rng.randint(0, 2)
            return g.seq(mn, rng.choice([None, mn, mn + 2]))
        if r < 0.9:
            return g.charf(rng.choice(["a", "ab", "bc"]))
        return tabs[rng.choice(names)]

    def alt(depth):
        r = rng.random()
        if r < 0.12:
            return g.empty ^ _const(rng.randint(-2, 2))
        k = rng.choice([1, 2, 2, 3])
        parts = [item(depth) for _ in range(k)]
        # at least one tabulation or terminal making progress: avoid pure empties
        p = parts[0]
        nargs = _nargs(parts[0])
        for q in parts[1:]:
            if rng.random() < 0.2:
                mn = rng.randint(0, 2)
                p = p.cat(q, hint=(mn, rng.choice([None, mn + 3]), rng.randint(0, 1), None))
            else:
                p = p * q
            nargs += _nargs(q)
        if rng.random() < 0.3:
            lo, hi = rng.randint(0, 3), rng.randint(3, 9)
            p = p.filter(A.cfun(lambda i, j, lo=lo, hi=hi: lo <= j - i <= hi, "i,j", "return j-i>=%d && j-i<=%d;" % (lo, hi)))
        if nargs == 0:
            return p ^ _const(rng.randint(-2, 2), 0)
        return p ^ mapf(rng, nargs)

    for nm in names:
        nalt = rng.randint(1, 4)
        alts = [alt(0) for _ in range(nalt)]
        # a base case that always exists keeps most cells valid
        if rng.random() < 0.7:
            alts.insert(rng.randint(0, len(alts)), g.el ^ mapf(rng, 1))

        def thunk(alts=alts):
            p = alts[0]
            for a in alts[1:]:
                p = p | a
            return p.aggregate(agg)
        tabs[nm].thunk = thunk
    g.axiom = tabs[names[0]]
    return g


def _nargs(p):
    n = p.node
    if "term" in n:
        return 0 if n["term"] == "empty" else 1
    return 1


def _const(v, nargs=0):
    return A.cfun(lambda *a, v=v: v, ",".join("a%d" % q for q in range(nargs)), "return %d;" % v)


def two_track(seed, agg=None):
    rng = random.Random(seed)
    h = rng.choice(["max", "min"])
    _mapfn_ = _mapfn
    if agg:
        h, _mapfn_ = agg, _mulfn
    g = A.Grammar(tracks=2, h=h, header=HEADER)
    names = ["M", "G"] if rng.random() < 0.5 else ["M"]
    tabs = {nm: g.tabulate(nm, None, always_valid=rng.random() < 0.3) for nm in names}

    def tab():
        return tabs[rng.choice(names)]

    def alt():
        r = rng.random()
        if r < 0.15:
            return g.empty ^ _const(rng.randint(-1, 2))
        if r < 0.25:
            return g.empty1 ^ _mapfn_(rng, 1)
        if r < 0.35:
            return g.empty2 ^ _mapfn_(rng, 1)
        if r < 0.5:
            return g.el1 << tab() >> g.el2 ^ _mapfn_(rng, 3)
        if r < 0.6:
            mn = rng.randint(1, 2)
            return g.seq(mn, rng.choice([None, mn + 2])) << tab() ^ _mapfn_(rng, 2)
        if r < 0.7:
            mn = rng.randint(1, 2)
            return tab() >> g.seq(mn, rng.choice([None, mn + 1])) ^ _mapfn_(rng, 2)
        if r < 0.8:
            return g.el1 << tab() ^ _mapfn_(rng, 2)
        if r < 0.9:
            return tab() >> g.el2 ^ _mapfn_(rng, 2)
        return g.eli << tab() >> g.eli ^ _mapfn_(rng, 3)

    for nm in names:
        alts = [alt() for _ in range(rng.randint(1, 4))]
        alts.insert(0, g.empty ^ _const(0))

        def thunk(alts=alts):
            p = alts[0]
            for a in alts[1:]:
                p = p | a
            return p.aggregate(h)
        tabs[nm].thunk = thunk
    g.axiom = tabs["M"]
    return g


def text(rng, n, alphabet="abc"):
    return "".join(rng.choice(alphabet) for _ in range(n))


def nested_track(seed):
    """One-track grammars with nested aggregations: an aggregate (max or min,
    over one or two alternatives) below a concatenation, mapped with the
    monotone algebra functions, so the DP answer is the best derivation under
    the nested aggregations' semantics (brute force checks it)."""
    rng = random.Random(seed)
    h = rng.choice(["max", "min"])
    g = A.Grammar(tracks=1, h=h, window=rng.choice([0, 0, 5]), header=HEADER)
    names = ["T0", "T1"][:rng.randint(1, 2)]
    tabs = {nm: g.tabulate(nm, None) for nm in names}

    def leaf():
        r = rng.random()
        if r < 0.5:
            return tabs[rng.choice(names)]
        return g.el if r < 0.75 else g.eli

    def nest():
        def one():
            k = rng.choice([2, 2, 3])
            parts = [leaf() for _ in range(k)]
            p = parts[0]
            for q in parts[1:]:
                p = p * q
            return p ^ _mapfn(rng, k)
        inner = one() if rng.random() < 0.6 else (one() | one())
        return inner.aggregate(rng.choice([h, "max", "min"]))

    def alt():
        parts = [leaf() for _ in range(rng.choice([1, 2, 2]))]
        parts[rng.randrange(len(parts))] = nest()
        p = parts[0]
        for q in parts[1:]:
            p = p * q
        return p ^ _mapfn(rng, len(parts))

    for nm in names:
        alts = [alt() for _ in range(rng.randint(1, 3))]
        alts.insert(rng.randint(0, len(alts)), g.el ^ _mapfn(rng, 1))

        def thunk(alts=alts):
            p = alts[0]
            for a in alts[1:]:
                p = p | a
            return p.aggregate(h)
        tabs[nm].thunk = thunk
    g.axiom = tabs[names[0]]
    return g
