"""The reference's example grammars (src/main/scala/v4/examples) written with
dpx.adp, each with the C bodies of its CodeGen algebra and Python callables
of the same meaning (the callables serve the CPU parity references only; the device runs
the C bodies).  Usage:

    import dpx, dpx.examples as ex
    ctx = dpx.Context()
    el = ctx.grammar(ex.elmamun("max"))
    el.parse("1+2*3*4+5")          # (81, True, trace)
"""
from . import adp as A


def nussinov():
    """Nussinov.scala:8-73 (Nussinov78 grammar, NussinovAlgebra), the
    grammar behind bench2/nu.c; uppercase ACGU input."""
    g = A.Grammar(tracks=1, h="max")
    pairs = {("A", "U"), ("U", "A"), ("G", "U"), ("U", "G"), ("C", "G"), ("G", "C")}
    nil = A.cfun(lambda: 0, "", "return 0;")
    left = A.cfun(lambda c, a: a, "c,a", "return a;")
    right = A.cfun(lambda a, c: a, "a,c", "return a;")
    pair = A.cfun(lambda l, a, r: a + 1, "l,a,r", "return a+1;")
    split = A.cfun(lambda l, r: l + r, "l,r", "return l+r;")
    bp = A.cfun(lambda i, j: i + 2 <= j and (chr(g.in1(i)), chr(g.in1(j - 1))) in pairs, "i,j",
                "if (i+2>j) return false; char a=_in1[i],b=_in1[j-1]; "
                "return (a=='A'&&b=='U') || (a=='U'&&b=='A') || (a=='G'&&b=='U') || (a=='U'&&b=='G') || "
                "(a=='C'&&b=='G') || (a=='G'&&b=='C');")
    s = g.tabulate("s", lambda: ((g.empty ^ nil) | (g.el * s ^ left) | (s * g.el ^ right) |
                                 ((g.el * s * g.el).filter(bp) ^ pair) | (s * s ^ split)).aggregate("max"),
                   always_valid=True)
    g.axiom = s
    return g


def elmamun(h):
    """ElMamun.scala: the bill grammar with the buyer (min) or seller (max) algebra."""
    g = A.Grammar(tracks=1, h=h)
    add = A.cfun(lambda l, c, r: l + r, "l,c,r", "return l + r;")
    mul = A.cfun(lambda l, c, r: l * r, "l,c,r", "return l * r;")
    b = g.tabulate("M", lambda: (g.digit | (b * g.charf("+") * b ^ add) | (b * g.charf("*") * b ^ mul)).aggregate(h))
    g.axiom = b
    return g


def brackets(window=8):
    """Brackets.scala (window 8): digits inside correctly nested brackets, summed."""
    g = A.Grammar(tracks=1, h="max", window=window)
    are = A.cfun(lambda i, j: j > i + 1 and g.in1(i) == 40 and g.in1(j - 1) == 41, "i,j",
                 "return j > i+1 && _in1[i] == '(' && _in1[j-1] == ')';")
    mid = A.cfun(lambda c1, i, c2: i, "c1,i,c2", "return i;")
    add = A.cfun(lambda x, y: x + y, "x,y", "return x+y;")
    zero = A.cfun(lambda: 0, "", "return 0;")
    m = g.tabulate("M", lambda: (g.digit | ((g.char * m * g.char).filter(are) ^ mid) | (m * m ^ add)).aggregate("max"))
    g.tabulate("N", lambda: ((g.empty ^ zero) | ((g.char * g.tabs["N"] * g.char).filter(are) ^ mid) |
                             (g.tabs["N"] * g.tabs["N"] ^ add)).aggregate("max"))  # unused: stripped like the reference
    g.axiom = m
    return g


def seqalign(algebra="sw"):
    """SeqAlign.scala: arbitrary gap cost, O(n^3) two-track grammar; the
    Smith-Waterman (+10/-3, open -3, extend -1, clamped at 0) or the
    Needleman-Wunsch (+4/-3, open -15, extend -1) algebra."""
    g = A.Grammar(tracks=2, h="max")
    if algebra == "sw":
        op, ex, ma, mi, clamp = -3, -1, 10, -3, True
    else:
        op, ex, ma, mi, clamp = -15, -1, 4, -3, False
    cl = (lambda v: max(0, v)) if clamp else (lambda v: v)
    cc = "return s>0?s:0;" if clamp else "return s;"
    start = A.cfun(lambda: 0, "", "return 0;")
    gap1 = A.cfun(lambda s, a: cl(a + op + ex * (s[1] - s[0] - 1)), "g,a",
                  "int s=a+(%d)+(%d)*(g._2-g._1-1); %s" % (op, ex, cc))
    gap2 = A.cfun(lambda a, s: cl(a + op + ex * (s[1] - s[0] - 1)), "a,g",
                  "int s=a+(%d)+(%d)*(g._2-g._1-1); %s" % (op, ex, cc))
    pair = A.cfun(lambda c1, a, c2: a + (ma if c1 == c2 else mi), "c1,a,c2", "return a+(c1==c2?%d:%d);" % (ma, mi))
    m = g.tabulate("M", lambda: ((g.empty ^ start) | (g.seq() << m ^ gap1) | (m >> g.seq() ^ gap2) |
                                 (g.el1 << m >> g.el2 ^ pair)).aggregate("max"), always_valid=True)
    g.axiom = m
    return g


def swat_affine(match=2, mismatch=-3, gap_open=-5, gap_ext=-2):
    """SWatAffine.scala with the scores as parameters (the C bodies of
    SWatAlgebra with open/ext/match/mismatch substituted): the grammar of
    bench/dp_swat/comp1.cu."""
    g = A.Grammar(tracks=2, h="max")
    start = A.cfun(lambda: 0, "", "return 0;")
    z = A.cfun(lambda p: 0, "p", "return 0;")
    o1 = A.cfun(lambda c, a: max(0, a + gap_open), "c1,a", "return a+(%d)>0 ? a+(%d) : 0;" % (gap_open, gap_open))
    o2 = A.cfun(lambda a, c: max(0, a + gap_open), "a,c2", "return a+(%d)>0 ? a+(%d) : 0;" % (gap_open, gap_open))
    e1 = A.cfun(lambda c, a: max(0, a + gap_ext), "c1,a", "return a+(%d)>0 ? a+(%d) : 0;" % (gap_ext, gap_ext))
    e2 = A.cfun(lambda a, c: max(0, a + gap_ext), "a,c2", "return a+(%d)>0 ? a+(%d) : 0;" % (gap_ext, gap_ext))
    pair = A.cfun(lambda c1, a, c2: a + (match if c1 == c2 else mismatch), "c1,a,c2",
                  "return a + (c1==c2 ? %d : %d);" % (match, mismatch))
    M = g.tabulate("M", lambda: ((g.empty ^ start) | (g.el1 << M >> g.el2 ^ pair) | g1 | g2).aggregate("max"),
                   always_valid=True)
    g1 = g.tabulate("g1", lambda: ((g.empty1 ^ z) | (g.el1 << M ^ o1) | (g.el1 << g1 ^ e1)).aggregate("max"),
                    always_valid=True)
    g2 = g.tabulate("g2", lambda: ((g.empty2 ^ z) | (M >> g.el2 ^ o2) | (g2 >> g.el2 ^ e2)).aggregate("max"),
                    always_valid=True)
    g.axiom = M
    return g


def matrix_chain():
    """MatrixMult2.scala:10-27: matrices (rows, cols) as T2ii, cost min."""
    g = A.Grammar(tracks=1, alphabet="T2ii", alphabet_fmt="ii", h="min")
    single = A.cfun(lambda m: 0, "i", "return 0;")
    mult = A.cfun(lambda i, l, k, r, j: _wrap(l + r + g.in1(i)[0] * g.in1(k)[0] * g.in1(j - 1)[1]), "i,l,k,r,j",
                  "return l+r + _in1[i]._1*_in1[k]._1*_in1[j-1]._2;")
    m = g.tabulate("M", lambda: ((g.el ^ single) | (g.emptyi * m * g.emptyi * m * g.emptyi ^ mult)).aggregate("min"),
                   always_valid=True)
    g.axiom = m
    return g


def _wrap(v):
    """int32 arithmetic as on the JVM and on the device (-fwrapv)."""
    return (v + (1 << 31)) % (1 << 32) - (1 << 31)


# librna's bp_index pairs (CG, GC, GU, UG, AU, UA) over the codes A=1 C=2 G=3 U=4
_PAIRS = {(2, 3), (3, 2), (3, 4), (4, 3), (1, 4), (4, 1)}


def zuker(algebra="mfe"):
    """Zuker.scala:110-146 (ZukerGrammar), the grammar behind bench2/zu.c,
    with one of its algebras:

    * "mfe": ZukerMFE (Zuker.scala:60-88).  The bodies call the librna
      functions; the device evaluates them over the context's parameter file
      (no Python callables: the built-in kernel and the reference's vectors
      are its checks);
    * "count": ZukerCount (Zuker.scala:28-44), h = sum: the number of
      derivations of the sequence (int32, wrapping as on the JVM), with
      Python callables for the oracle."""
    count = algebra == "count"
    if algebra not in ("mfe", "count"):
        raise ValueError("algebra is mfe or count")
    h = "sum" if count else "min"
    g = A.Grammar(tracks=1, h=h, empty_value=None if count else "999999", rna=True)
    if count:
        sadd = A.cfun(lambda lb, e: e, "lb,e", "return e;")
        cadd = A.cfun(lambda x, e: _wrap(x * e), "x,e", "return x*e;")
        dlr = A.cfun(lambda lb, e, rb: e, "lb,e,rb", "return e;")
        sr = A.cfun(lambda lb, e, rb: e, "lb,e,rb", "return e;")
        hl = A.cfun(lambda lb, f1, x, f2, rb: 1, "lb,f1,x,f2,rb", "return 1;")
        bl = A.cfun(lambda lb, f1, b, x, f2, rb: x, "lb,f1,b,x,f2,rb", "return x;")
        br = A.cfun(lambda lb, f1, x, b, f2, rb: x, "lb,f1,x,b,f2,rb", "return x;")
        il = A.cfun(lambda f1, f2, r1, x, r2, f3, f4: x, "f1,f2,r1,x,r2,f3,f4", "return x;")
        ml = A.cfun(lambda lb, f1, x, f2, rb: x, "lb,f1,x,f2,rb", "return x;")
        app = A.cfun(lambda c1, c: _wrap(c1 * c), "c1,c", "return c1*c;")
        ul = A.cfun(lambda c1: c1, "c1", "return c1;")
        addss = A.cfun(lambda c1, e: c1, "c1,e", "return c1;")
        nil = A.cfun(lambda: 1, "", "return 1;")
    else:
        f = lambda args, body: A.cfun(None, args, body)
        sadd = f("lb,e", "return e;")
        cadd = f("x,e", "return x+e;")
        dlr = f("lb,e,rb", "return e + ext_mismatch_energy(lb,rb-1) + termau_energy(lb,rb-1);")
        sr = f("lb,e,rb", "return e + sr_energy(lb,rb);")
        hl = f("lb,f1,x,f2,rb", "return hl_energy(f1,f2) + sr_energy(lb,rb);")
        bl = f("lb,f1,b,x,f2,rb", "return x + bl_energy(f1,b._1,b._2-1,f2,f2-1) + sr_energy(lb,rb);")
        br = f("lb,f1,x,b,f2,rb", "return x + br_energy(f1,b._1,b._2-1,f2,f1+1) + sr_energy(lb,rb);")
        il = f("f1,f2,r1,x,r2,f3,f4", "return x + il_energy(f2,r1._2,r2._1-1,f3) + sr_energy(f1,f4);")
        ml = f("lb,f1,x,f2,rb", "return ml_energy() + ul_energy() + x + termau_energy(f1,f2) + sr_energy(lb,rb) + "
                                "ml_mismatch_energy(f1,f2);")
        app = f("c1,c", "return c1+c;")
        ul = f("c1", "return ul_energy()+c1;")
        addss = f("c1,e", "return c1;")
        nil = f("", "return 0;")
    tail = A.cfun(lambda i, j: j == len(g._in1), "i,j", "return j==M_W-1;")
    stackpairing = A.cfun(lambda i, j: i + 4 <= j and (g.in1(i), g.in1(j - 1)) in _PAIRS and
                          (g.in1(i + 1), g.in1(j - 2)) in _PAIRS, "i,j",
                          "return i+4<=j && bp_index(_in1[i],_in1[j-1])!=NO_BP && bp_index(_in1[i+1],_in1[j-2])!=NO_BP;")
    BASE, LOC, REG, REG3, REG30 = g.eli, g.emptyi, g.seq(), g.seq(3), g.seq(1, 30)
    st = g.tabulate("st", lambda: ((BASE * st ^ sadd) | (dangle() * st ^ cadd) | (g.empty ^ nil)).aggregate(h).filter(tail),
                    always_valid=True)
    cl = g.tabulate("cl", lambda: (
        (BASE * cl * BASE ^ sr) |
        (BASE * BASE * REG3 * BASE * BASE ^ hl) |
        (BASE * BASE * REG30 * cl * BASE * BASE ^ bl).aggregate(h) |
        (BASE * BASE * cl * REG30 * BASE * BASE ^ br).aggregate(h) |
        (BASE * BASE * REG30 * cl * REG30 * BASE * BASE ^ il).aggregate(h) |
        (BASE * BASE * mlc * BASE * BASE ^ ml)).aggregate(h).filter(stackpairing))

    def dangle():
        return LOC * cl * LOC ^ dlr

    mlc = g.tabulate("ml", lambda: ((BASE * mlc ^ sadd) | ((dangle() ^ ul) * ml1 ^ app)).aggregate(h))
    ml1 = g.tabulate("ml1", lambda: ((BASE * ml1 ^ sadd) | ((dangle() ^ ul) * ml1 ^ app) | (dangle() ^ ul) |
                                     ((dangle() ^ ul) * REG ^ addss)).aggregate(h))
    g.axiom = st
    return g


def rnafold():
    """RNAFold.scala:10-66 (RNAFoldGrammar + RNAFoldAlgebra), librna calls
    evaluated on the device."""
    g = A.Grammar(tracks=1, h="min", rna=True)
    f = lambda args, body: A.cfun(None, args, body)
    hairpin = f("ij", "return hl_energy(ij._1,ij._2-1);")
    stack = f("i,s,j", "return sr_energy(i,j) + s;")
    iloop = f("ik,s,lj", "return il_energy(ik._1,ik._2,lj._1-1,lj._2-1) + s;")
    mloop = f("i,s,j", "return s;")
    left = f("l,r", "return l;")
    right = f("l,r", "return r;")
    join = f("l,r", "return l+r;")
    basepairing = f("i,j", "return i+2<=j && bp_index(_in1[i],_in1[j-1])!=NO_BP;")
    length = lambda mn: f("i,j", "return j-i>=%d;" % mn)
    Qp = g.tabulate("Qp", lambda: ((g.seq(3) ^ hairpin) | (g.eli * Qp * g.eli ^ stack) |
                                   (g.seq(1, 30) * Qp * g.seq(1, 30) ^ iloop) |
                                   (g.eli * QM * g.eli ^ mloop)).aggregate("min").filter(basepairing))
    QM = g.tabulate("QM", lambda: (Q * Q ^ join).filter(length(4)).aggregate("min"))
    Q = g.tabulate("Q", lambda: (QM | (Q * g.eli ^ left) | (g.eli * Q ^ right) | Qp).filter(length(2)).aggregate("min"))
    g.axiom = Q
    return g

