"""Example grammars (dpx.examples) plus the oracle helpers of the tests."""
from dpx.examples import *  # noqa: F401,F403
from dpx.examples import _wrap  # noqa: F401


def oracle(g):
    import adp_oracle
    return adp_oracle.Oracle(g.spec(), {n: f.py for n, f in g.functions().items()})


def run_oracle(g, in1, in2=None, trace=True):
    o = oracle(g)
    g._in1 = g.py_input(in1)
    g._in2 = g.py_input(in2) if in2 is not None else None
    return o, o.run(g._in1, g._in2, trace=trace)


def _wrapper(fmt):
    """Integer arithmetic of a struct format as on the device (-fwrapv)."""
    import struct
    bits = 8 * struct.calcsize("=" + fmt)
    if fmt in "fd":
        return lambda v: v
    return lambda v: (v + (1 << (bits - 1))) % (1 << bits) - (1 << (bits - 1))


def motzkin(answer="int", fmt="i", h="sum", window=0):
    """s -> empty | el s | el s el s with the counting algebra (h = sum): the
    number of derivations of s over n letters is the Motzkin number M_n
    (secondary structures of n bases with any pairing, OEIS A001006)."""
    import dpx.adp as A
    w = _wrapper(fmt)
    g = A.Grammar(tracks=1, answer=answer, answer_fmt=fmt, h=h, window=window)
    one = A.cfun(lambda: 1, "", "return 1;")
    skip = A.cfun(lambda c, x: x, "c,x", "return x;")
    pair = A.cfun(lambda a, x, b, y: w(x * y), "a,x,b,y", "return x*y;")
    s = g.tabulate("s", lambda: ((g.empty ^ one) | (g.el * s ^ skip) | (g.el * s * g.el * s ^ pair)).aggregate(h))
    g.axiom = s
    return g


def splits(h="count", window=0):
    """s -> el s s | el: with h = count a cell's answer is the number of
    split points whose two sides parse (one for a single letter)."""
    import dpx.adp as A
    g = A.Grammar(tracks=1, h=h, window=window)
    f3 = A.cfun(lambda c, x, y: x + y, "c,x,y", "return x+y;")
    f1 = A.cfun(lambda c: 1, "c", "return 1;")
    s = g.tabulate("s", lambda: ((g.el * s * s ^ f3) | (g.el ^ f1)).aggregate(h))
    g.axiom = s
    return g


def motzkin_numbers(n, bits=None):
    """M_0..M_n by M_L = M_{L-1} + sum_k M_k M_{L-2-k} (exact integers,
    reduced to signed `bits` when given)."""
    m = [1]
    for L in range(1, n + 1):
        m.append(m[L - 1] + sum(m[k] * m[L - 2 - k] for k in range(L - 1)))
    if bits:
        m = [(v + (1 << (bits - 1))) % (1 << bits) - (1 << (bits - 1)) for v in m]
    return m


def split_counts(n):
    """(valid, count) of splits() over lengths 0..n, from the grammar by hand:
    a length L >= 3 cell has the splits a + b = L - 1 with both sides valid."""
    valid, count = [False, True], [0, 1]
    for L in range(2, n + 1):
        c = sum(1 for a in range(1, L - 1) if valid[a] and valid[L - 1 - a])
        valid.append(c > 0)
        count.append(c)
    return valid[:n + 1], count[:n + 1]


def edit_distance():
    """E -> empty | el1 E | E el2 | el1 E el2 (two tracks, h = min, with the
    reference's empty value): the Levenshtein distance of the two inputs."""
    import dpx.adp as A
    g = A.Grammar(tracks=2, h="min", empty_value="1000000")
    zero = A.cfun(lambda: 0, "", "return 0;")
    dl = A.cfun(lambda c, a: a + 1, "c,a", "return a+1;")
    ins = A.cfun(lambda a, c: a + 1, "a,c", "return a+1;")
    sub = A.cfun(lambda c1, a, c2: a + (0 if c1 == c2 else 1), "c1,a,c2", "return a+(c1==c2?0:1);")
    E = g.tabulate("E", lambda: ((g.empty ^ zero) | (g.el1 << E ^ dl) | (E >> g.el2 ^ ins) |
                                 (g.el1 << E >> g.el2 ^ sub)).aggregate("min"))
    g.axiom = E
    return g


def levenshtein(a, b):
    """The textbook edit-distance recurrence, row by row."""
    prev = list(range(len(b) + 1))
    for x in range(1, len(a) + 1):
        cur = [x] + [0] * len(b)
        for y in range(1, len(b) + 1):
            cur[y] = min(prev[y] + 1, cur[y - 1] + 1, prev[y - 1] + (a[x - 1] != b[y - 1]))
        prev = cur
    return prev[len(b)]


def nested_pairs():
    """Nussinov with the closed pair's interior as a nested aggregation
    (s -> empty | el s | (el (s s ^ split aggregate max) el filter bp) ^ pair),
    the shape of the reference's aggregated loops (Zuker2.scala:111-112,
    MatrixMult3.scala:41-43)."""
    import dpx.adp as A
    g = A.Grammar(tracks=1, h="max")
    pairs = {("A", "U"), ("U", "A"), ("G", "C"), ("C", "G"), ("G", "U"), ("U", "G")}
    nil = A.cfun(lambda: 0, "", "return 0;")
    left = A.cfun(lambda c, a: a, "c,a", "return a;")
    pair = A.cfun(lambda l, a, r: a + 1, "l,a,r", "return a+1;")
    split = A.cfun(lambda l, r: l + r, "l,r", "return l+r;")
    bp = A.cfun(lambda i, j: i + 2 <= j and (chr(g.in1(i)), chr(g.in1(j - 1))) in pairs, "i,j",
                "if (i+2>j) return false; char a=_in1[i],b=_in1[j-1]; "
                "return (a=='A'&&b=='U')||(a=='U'&&b=='A')||(a=='G'&&b=='C')||(a=='C'&&b=='G')||"
                "(a=='G'&&b=='U')||(a=='U'&&b=='G');")
    s = g.tabulate("s", lambda: ((g.empty ^ nil) | (g.el * s ^ left) |
                                 ((g.el * (s * s ^ split).aggregate("max") * g.el).filter(bp) ^ pair)).aggregate("max"),
                   always_valid=True)
    g.axiom = s
    return g


def left_terms(wide):
    """A Nussinov-shaped grammar whose split rule maps its left operand
    before combining it: s -> empty | el s | (s ^ widen) s ^ comb.  widen is
    a left term of the middle product (a call of the staged left operand
    alone); it returns a long long (wide) or an int, so the generated code
    keeps it in LDS only in the int case and evaluates it in place
    otherwise.  comb(widen(a), b) = a + b + 1: monotone, with many ties."""
    g = A.Grammar(tracks=1, h="max")
    nil = A.cfun(lambda: 0, "", "return 0;")
    left = A.cfun(lambda c, a: a, "c,a", "return a;")
    widen = A.cfun(lambda a: 3 * a + 1, "a", "return (long long)a * 3 + 1;" if wide else "return a * 3 + 1;")
    comb = A.cfun(lambda x, y: (x - 1) // 3 + y + 1, "x,y", "return (int)((x - 1) / 3) + y + 1;")
    # both sides non-empty, as Concat.indices gives s ~ s (BaseParsers.scala:279-283)
    s = g.tabulate("s", lambda: ((g.empty ^ nil) | (g.el * s ^ left) |
                                 ((s ^ widen).cat(s, hint=(1, None, 1, None)) ^ comb)).aggregate("max"),
                   always_valid=True)
    g.axiom = s
    return g
