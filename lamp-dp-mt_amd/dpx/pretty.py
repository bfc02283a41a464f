"""Replay of backtraces through the reference's pretty-print algebras.

The reference returns a backtrace (a list of (cell, rule, split positions)
records in BFS order from the axiom) and `Tabulate.build` replays it with
another algebra of the same signature, typically a pretty printer
(BaseParsers.scala:189-192).  These functions do the same for the records the
C ABI returns (dpx_trace_t as tuples (i, j, rule, pos0, pos1)): every record
is evaluated after the records it refers to (reverse BFS order), with the
pretty algebra of its grammar:

  zuker(seq, trace)        -> dot-bracket string   (ZukerPrettyPrint, Zuker.scala:46-62)
  nussinov(seq, trace)     -> dot-bracket string   (NussinovString, Nussinov.scala:27-34)
  matchain(rows, cols, tr) -> "((|10x100|*|100x5|)*|5x50|)"  (MatrixPrettyPrint, MatrixMult.scala:39-44)
  swat(a, b, trace)        -> (track 1, track 2)   (SWatPretty, SWatAffine.scala:36-47)
  rnafold(seq, trace)      -> dot-bracket string   (RNAFoldPrettyPrint, RNAFold.scala:31-41)

Rule ids are the reference's (bench2/zu.h:2-5, bench2/nu.h, dp_swat/comp1.h).
"""


def _dots(k):
    return "." * max(0, k)


def zuker(seq, trace):
    """Dot-bracket structure of a Zuker MFE backtrace.

    Tables by rule id: cl 0-5, ml 6-7, ml1 8-11, st 12-14.  Productions
    (Zuker.scala:64-90 with the pretty algebra): sr "("+cl(i+1,j-1)+")",
    hl "(("+dots+"))", bl "(("+dots+cl(k,j-2)+"))", br "(("+cl(i+2,k)+dots+"))",
    il "(("+dots+cl(l,k)+dots+"))", ml "(("+ml(i+2,j-2)+"))"; sadd "."+x;
    cadd / app concatenate; dlr and ul add nothing; addss appends dots; nil "".
    """
    ans = {}

    def get(tab, i, j):
        return ans[(tab, i, j)]

    for (i, j, rule, p0, p1) in reversed(trace):
        if rule == 0:
            v, tab = "(" + get("cl", i + 1, j - 1) + ")", "cl"
        elif rule == 1:
            v, tab = "((" + _dots(j - i - 4) + "))", "cl"
        elif rule == 2:
            v, tab = "((" + _dots(p0 - i - 2) + get("cl", p0, j - 2) + "))", "cl"
        elif rule == 3:
            v, tab = "((" + get("cl", i + 2, p0) + _dots(j - 2 - p0) + "))", "cl"
        elif rule == 4:
            v, tab = "((" + _dots(p0 - i - 2) + get("cl", p0, p1) + _dots(j - 2 - p1) + "))", "cl"
        elif rule == 5:
            v, tab = "((" + get("ml", i + 2, j - 2) + "))", "cl"
        elif rule == 6:
            v, tab = "." + get("ml", i + 1, j), "ml"
        elif rule == 7:
            v, tab = get("cl", i, p0) + get("ml1", p0, j), "ml"
        elif rule == 8:
            v, tab = "." + get("ml1", i + 1, j), "ml1"
        elif rule == 9:
            v, tab = get("cl", i, p0) + get("ml1", p0, j), "ml1"
        elif rule == 10:
            v, tab = get("cl", i, j), "ml1"
        elif rule == 11:
            v, tab = get("cl", i, p0) + _dots(j - p0), "ml1"
        elif rule == 12:
            v, tab = "." + get("st", i + 1, j), "st"
        elif rule == 13:
            v, tab = get("cl", i, p0) + get("st", p0, j), "st"
        elif rule == 14:
            v, tab = "", "st"
        else:
            raise ValueError("not a Zuker rule id: %r" % (rule,))
        ans[(tab, i, j)] = v
    i, j = trace[0][0], trace[0][1]
    return ans[("st", i, j)]


def rnafold(seq, trace):
    """Dot-bracket structure of an RNAfold backtrace (None for the empty answer).

    Rule ids: Q 0 QM, 1 left, 2 right, 3 Qp; QM 4 join(pos0); Qp 5 hairpin,
    6 stack, 7 iloop(pos0 = k, pos1 = l), 8 mloop.  Productions of the pretty
    algebra (RNAFold.scala:31-41, grammar :43-64): hairpin "("+dots+")",
    stack and mloop "("+x+")", iloop "("+dots(k-1-i)+Qp(k,l)+dots(j-1-l)+")",
    left x+".", right "."+x, join x+y.
    """
    if not trace:
        return None
    ans = {}
    for (i, j, rule, p0, p1) in reversed(trace):
        if rule == 0:
            v, tab = ans[("QM", i, j)], "Q"
        elif rule == 1:
            v, tab = ans[("Q", i, j - 1)] + ".", "Q"
        elif rule == 2:
            v, tab = "." + ans[("Q", i + 1, j)], "Q"
        elif rule == 3:
            v, tab = ans[("Qp", i, j)], "Q"
        elif rule == 4:
            v, tab = ans[("Q", i, p0)] + ans[("Q", p0, j)], "QM"
        elif rule == 5:
            v, tab = "(" + _dots(j - i - 2) + ")", "Qp"
        elif rule == 6:
            v, tab = "(" + ans[("Qp", i + 1, j - 1)] + ")", "Qp"
        elif rule == 7:
            v, tab = "(" + _dots(p0 - 1 - i) + ans[("Qp", p0, p1)] + _dots(j - 1 - p1) + ")", "Qp"
        elif rule == 8:
            v, tab = "(" + ans[("QM", i + 1, j - 1)] + ")", "Qp"
        else:
            raise ValueError("not an RNAfold rule id: %r" % (rule,))
        ans[(tab, i, j)] = v
    return ans[("Q", trace[0][0], trace[0][1])]


def nussinov(seq, trace):
    """Dot-bracket structure of a Nussinov backtrace (rules 0 nil, 1 left,
    2 right, 3 pair, 4 split at pos0)."""
    ans = {}
    for (i, j, rule, p0, _p1) in reversed(trace):
        if rule == 0:
            v = ""
        elif rule == 1:
            v = "." + ans[(i + 1, j)]
        elif rule == 2:
            v = ans[(i, j - 1)] + "."
        elif rule == 3:
            v = "(" + ans[(i + 1, j - 1)] + ")"
        elif rule == 4:
            v = ans[(i, p0)] + ans[(p0, j)]
        else:
            raise ValueError("not a Nussinov rule id: %r" % (rule,))
        ans[(i, j)] = v
    return ans[(trace[0][0], trace[0][1])]


def matchain(rows, cols, trace):
    """Parenthesisation of a matrix-chain backtrace (rule 0 single, 1 mult at pos0)."""
    ans = {}
    for (i, j, rule, p0, _p1) in reversed(trace):
        if rule == 0:
            v = "|%dx%d|" % (rows[i], cols[i])
        elif rule == 1:
            v = "(" + ans[(i, p0)] + "*" + ans[(p0, j)] + ")"
        else:
            raise ValueError("not a matrix-chain rule id: %r" % (rule,))
        ans[(i, j)] = v
    return ans[(trace[0][0], trace[0][1])]


def swat(a, b, trace):
    """The two alignment tracks of a Smith-Waterman/Gotoh backtrace.

    Tables: M rules 0-3 (start, pair, g1, g2), g1 rules 4-6 (zero, open, extend
    in track 1), g2 rules 7-9 (zero, open, extend in track 2); prefix lengths
    (i, j) index a and b (SWatAffine.scala:50-69).
    """
    ans = {}
    for (i, j, rule, _p0, _p1) in reversed(trace):
        if rule == 0:
            v, tab = (".", "."), "M"
        elif rule == 1:
            x = ans[("M", i - 1, j - 1)]
            v, tab = (x[0] + a[i - 1], x[1] + b[j - 1]), "M"
        elif rule == 2:
            v, tab = ans[("g1", i, j)], "M"
        elif rule == 3:
            v, tab = ans[("g2", i, j)], "M"
        elif rule == 4:  # empty1 at i == 0 yields j (TTParsers.scala:39-41)
            v, tab = (".", "." + "-" * j), "g1"
        elif rule in (5, 6):
            x = ans[("M" if rule == 5 else "g1", i - 1, j)]
            v, tab = (x[0] + a[i - 1], x[1] + "-"), "g1"
        elif rule == 7:  # empty2 at j == 0 yields i (TTParsers.scala:42-44)
            v, tab = ("." + "-" * i, "."), "g2"
        elif rule in (8, 9):
            x = ans[("M" if rule == 8 else "g2", i, j - 1)]
            v, tab = (x[0] + "-", x[1] + b[j - 1]), "g2"
        else:
            raise ValueError("not a SW rule id: %r" % (rule,))
        ans[(tab, i, j)] = v
    return ans[("M", trace[0][0], trace[0][1])]
