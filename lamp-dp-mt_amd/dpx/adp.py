"""ADP grammar combinators for the generic path (dpx_grammar_*).

Mirrors the reference's parser combinators (src/main/scala/v4/BaseParsers.scala,
ADPParsers.scala, TTParsers.scala, LexicalParsers in ADPParsers.scala:60-86)
with Python operators, and produces the JSON grammar tree that
dpx_grammar_create compiles to HIP:

    Scala                      here
    p1 | p2                    p1 | p2
    p1 ~ p2                    p1 * p2
    p1 -~ p2   (two tracks)    p1 << p2
    p1 ~- p2   (two tracks)    p1 >> p2
    p ^^ f                     p ^ f
    p filter f                 p.filter(f)
    p aggregate h              p.aggregate(h)
    p1 ~ (a,b,c,d) ~ p2        p1.cat(p2, hint=(a, b, c, d))
    tabulate("M", p, true)     g.tabulate("M", lambda: p, always_valid=True)
    cfun2(fn, "a,b", body)     cfun(fn, "a,b", body)

Python's precedence (* and << >> above ^ above |) matches the Scala use:
`el * s * el ^ pair | s * s ^ split` reads ((el*s*el) ^ pair) | ((s*s) ^ split).
Each function carries its C body (compiled for the device) and a Python
callable of the same meaning (used by the CPU oracle in the tests); char
inputs reach the Python callables as ints (byte values).
"""
import itertools
import json
import struct

_ids = itertools.count()


class Fn:
    """An algebra function: Python callable + argument names + C body
    (the reference's CFun, BaseParsers.scala:13-27)."""

    def __init__(self, py, args, body, name=None):
        self.py = py
        self.args = args
        self.body = body
        self.name = name or "fn%d" % next(_ids)


def cfun(py, args, body, name=None):
    return Fn(py, args, body, name)


class P:
    def __init__(self, g, node, fns=()):
        self.g = g
        self.node = node
        self.fns = list(fns)

    def _bin(self, other, node):
        return P(self.g, node, self.fns + other.fns)

    def __or__(self, o):
        return self._bin(o, {"or": [self.node, o.node]})

    def __mul__(self, o):
        return self._bin(o, {"cat": [self.node, o.node], "track": 0})

    def cat(self, o, hint=None):
        n = {"cat": [self.node, o.node], "track": 0}
        if hint is not None:
            n["hint"] = [(-1 if x is None else x) for x in hint]
        return self._bin(o, n)

    def __lshift__(self, o):
        return self._bin(o, {"cat": [self.node, o.node], "track": 1})

    def __rshift__(self, o):
        return self._bin(o, {"cat": [self.node, o.node], "track": 2})

    def __xor__(self, f):
        return P(self.g, {"map": self.node, "f": f.name}, self.fns + [f])

    def filter(self, f):
        return P(self.g, {"filter": self.node, "f": f.name}, self.fns + [f])

    def aggregate(self, h):
        node, fns = _agg(h)
        return P(self.g, {"agg": self.node, "h": node}, self.fns + fns)


def _agg(h):
    if isinstance(h, str):
        return h, []
    kind, f = h
    return {kind: f.name}, [f]


def max_by(f):
    return ("maxBy", f)


def min_by(f):
    return ("minBy", f)


class Tab(P):
    def __init__(self, g, name, thunk, always_valid):
        super().__init__(g, {"tab": name})
        self.name = name
        self.thunk = thunk
        self.always_valid = always_valid
        self._p = None

    def inner(self):
        if self._p is None:
            self._p = self.thunk()
        return self._p


class Grammar:
    """A grammar with its signature types.  tracks=1 for ADPParsers (one
    sequence, subwords (i, j)), 2 for TTParsers (two sequences, cells
    (i, j) = prefixes).  alphabet_fmt / answer_fmt are struct formats of one
    input element / one answer as the C types lay them out."""

    def __init__(self, tracks=1, alphabet="char", alphabet_fmt="b", answer="int", answer_fmt="i", h="max",
                 window=0, empty_value=None, header="", rna=False):
        self.tracks = tracks
        self.alphabet, self.alphabet_fmt = alphabet, alphabet_fmt
        self.answer, self.answer_fmt = answer, answer_fmt
        self.h = h
        self.window = window
        self.empty_value = empty_value  # the reference's cudaEmpty (CodeGen.scala:519)
        self.header = header
        self.rna = rna  # bodies call librna; inputs are RNA letters, passed as codes A=1 C=2 G=3 U=4
        self.tabs = {}
        self.axiom = None
        self._in1 = self._in2 = None
        t = lambda name, **kw: P(self, dict({"term": name}, **kw))
        self.empty = t("empty")
        if tracks == 1:
            self.emptyi, self.el, self.eli = t("emptyi"), t("el"), t("eli")
        else:
            self.empty1, self.empty2, self.el1, self.el2, self.eli = t("empty1"), t("empty2"), t("el1"), t("el2"), t("eli")

    def seq(self, min=1, max=None):
        n = {"term": "seq", "min": min}
        if max is not None:
            n["max"] = max
        return P(self, n)

    def tabulate(self, name, thunk, always_valid=False):
        if name in self.tabs:
            raise ValueError("duplicate tabulation name " + name)
        t = Tab(self, name, thunk, always_valid)
        self.tabs[name] = t
        return t

    # inputs for the Python callables (the reference's in(k) / in1(k) / in2(k))
    def in1(self, k):
        return self._in1[k]

    def in2(self, k):
        return self._in2[k]

    # ---- LexicalParsers (ADPParsers.scala:60-86) ------------------------
    def _char_filter(self, py, cbody):
        return cfun(lambda i, j: i + 1 == j and py(self.in1(i)), "i,j",
                    "if (i + 1 != j) return false; const int c = _in1[i]; " + cbody)

    @property
    def char(self):
        return self.el

    def charf(self, chars):
        cs = "".join("'%s'," % (c if c not in "\\'" else "\\" + c) for c in chars)
        return self.el.filter(self._char_filter(lambda c: chr(c) in chars,
                                                "const char s[] = {%s 0}; for (int q = 0; s[q]; ++q) if (s[q] == c) return true; return false;" % cs))

    @property
    def digit(self):
        if not hasattr(self, "_digit"):
            isd = self._char_filter(lambda c: 48 <= c <= 57, "return c >= '0' && c <= '9';")
            val = cfun(lambda c: c - 48, "c", "return c - '0';")
            self._digit = self.el.filter(isd) ^ val
        return self._digit

    @property
    def letter(self):
        return self.el.filter(self._char_filter(lambda c: 65 <= c <= 90 or 97 <= c <= 122,
                                                "return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z');"))

    # ---- spec --------------------------------------------------------------
    def functions(self):
        fns = {}
        todo = list(self.tabs.values())
        for t in todo:
            for f in t.inner().fns:
                fns[f.name] = f
        if not isinstance(self.h, str):
            fns[self.h[1].name] = self.h[1]
        return fns

    def spec(self):
        if self.axiom is None:
            raise ValueError("set grammar.axiom")
        fns = self.functions()
        h, _ = _agg(self.h)
        d = {"tracks": self.tracks, "alphabet": self.alphabet, "alphabet_bytes": struct.calcsize("=" + self.alphabet_fmt),
             "answer": self.answer, "answer_bytes": struct.calcsize("=" + self.answer_fmt), "header": self.header,
             "h": h, "window": self.window, "axiom": self.axiom.name,
             "functions": {n: {"args": f.args, "body": f.body} for n, f in fns.items()},
             "tabulations": {n: {"always_valid": t.always_valid, "p": t.inner().node} for n, t in self.tabs.items()}}
        if self.empty_value is not None:
            d["empty"] = self.empty_value
        if self.rna:
            d["rna"] = True
        return d

    def spec_json(self):
        return json.dumps(self.spec())

    def pack(self, seq):
        """Input elements as the device reads them."""
        if isinstance(seq, (bytes, bytearray)):
            return bytes(seq)
        if isinstance(seq, str):
            if self.rna:  # RNASignature.convert (the reference maps other letters to 5, past its tables)
                codes = {"A": 1, "C": 2, "G": 3, "U": 4, "T": 4}
                try:
                    return bytes(codes[c] for c in seq.upper())
                except KeyError:
                    raise ValueError("RNA input must be ACGU/T")
            return seq.encode()
        fmt = "=" + self.alphabet_fmt
        return b"".join(struct.pack(fmt, *(x if isinstance(x, tuple) else (x,))) for x in seq)

    def py_input(self, seq):
        """Input elements as the Python callables see them (RNA letters as
        the codes the device reads)."""
        if isinstance(seq, str) and self.rna:
            return list(self.pack(seq))
        if isinstance(seq, (bytes, bytearray, str)):
            return list(seq.encode() if isinstance(seq, str) else seq)
        return list(seq)

    def unpack_answer(self, raw):
        v = struct.unpack("=" + self.answer_fmt, raw[:struct.calcsize("=" + self.answer_fmt)])
        return v[0] if len(v) == 1 else v
