"""Python view of libdpx.so, the MI355X DP backend (C ABI in include/dpx.h).

The library is the product; this module only marshals arguments.  It fails
loudly when libdpx.so is missing: there is no CPU fallback.
"""
import ctypes
import os

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "libdpx.so")
PAR_TURNER2004 = os.path.join(PKG, "params", "rna_turner2004.par")

OK, EINVAL, ENOMEM, EHIP, ENOPARAMS, EPARAMS, ETRACE, EBUSY, EAGAIN = range(9)

# exported symbols, as declared in include/dpx.h
EXPORTS = [
    "dpx_create", "dpx_destroy", "dpx_last_error", "dpx_set_rna_params",
    "dpx_nussinov", "dpx_zuker", "dpx_swat", "dpx_matchain",
    "dpx_zuker_batch", "dpx_nussinov_batch", "dpx_last_timing",
]


class DpxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("dpx error %d: %s" % (code, msg))
        self.code = code


class Trace(ctypes.Structure):
    _fields_ = [("i", ctypes.c_int32), ("j", ctypes.c_int32), ("rule", ctypes.c_int32),
                ("pos", ctypes.c_int32 * 2)]


class SwatScores(ctypes.Structure):
    _fields_ = [("match", ctypes.c_int32), ("mismatch", ctypes.c_int32),
                ("gap_open", ctypes.c_int32), ("gap_ext", ctypes.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libdpx.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.dpx_last_error.restype = ctypes.c_char_p
        L.dpx_last_error.argtypes = [ctypes.c_void_p]
        L.dpx_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]
        L.dpx_destroy.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _tuples(buf, n):
    return [(t.i, t.j, t.rule, t.pos[0], t.pos[1]) for t in buf[:n]]


class Context:
    """One device context: resident energy tables and DP scratch."""

    def __init__(self, device=0, rna_params=PAR_TURNER2004):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.dpx_create(ctypes.byref(h), device)
        if rc != OK:
            raise DpxError(rc, "dpx_create(%d) failed" % device)
        self.h = h
        if rna_params:
            self.set_rna_params(rna_params)

    def close(self):
        if self.h:
            self._L.dpx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            raise DpxError(rc, self._L.dpx_last_error(self.h).decode())

    def set_rna_params(self, path):
        self._check(self._L.dpx_set_rna_params(self.h, path.encode()))

    # ---- single instances -------------------------------------------------
    def nussinov(self, seq, min_loop=0, trace=True):
        n = len(seq)
        cap = 4 * n + 8
        tr = (Trace * cap)() if trace else None
        score, tl = ctypes.c_int32(), ctypes.c_int()
        self._check(self._L.dpx_nussinov(self.h, seq.encode(), n, min_loop, ctypes.byref(score),
                                         tr, cap, ctypes.byref(tl)))
        return score.value, (_tuples(tr, tl.value) if trace else None)

    def zuker(self, seq, trace=True):
        n = len(seq)
        cap = 4 * n + 8
        tr = (Trace * cap)() if trace else None
        score, tl = ctypes.c_int32(), ctypes.c_int()
        self._check(self._L.dpx_zuker(self.h, seq.encode(), n, ctypes.byref(score), tr, cap, ctypes.byref(tl)))
        return score.value, (_tuples(tr, tl.value) if trace else None)

    def matchain(self, rows, cols, trace=True):
        n = len(rows)
        cap = 2 * n + 8
        tr = (Trace * cap)() if trace else None
        r = (ctypes.c_int32 * n)(*rows)
        c = (ctypes.c_int32 * n)(*cols)
        score, tl = ctypes.c_int32(), ctypes.c_int()
        self._check(self._L.dpx_matchain(self.h, r, c, n, ctypes.byref(score), tr, cap, ctypes.byref(tl)))
        return score.value, (_tuples(tr, tl.value) if trace else None)

    def swat(self, a, b, match, mismatch, gap_open, gap_ext, trace_from=0, trace=True):
        n1, n2 = len(a), len(b)
        cap = 2 * (n1 + n2) + 8
        tr = (Trace * cap)() if trace else None
        sc = SwatScores(match, mismatch, gap_open, gap_ext)
        score, best, bi, bj, tl = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int()
        self._check(self._L.dpx_swat(self.h, a.encode(), n1, b.encode(), n2, ctypes.byref(sc), ctypes.byref(score),
                                     ctypes.byref(best), ctypes.byref(bi), ctypes.byref(bj), trace_from,
                                     tr, cap, ctypes.byref(tl)))
        return score.value, (best.value, bi.value, bj.value), (_tuples(tr, tl.value) if trace else None)

    # ---- batches ----------------------------------------------------------
    def _batch(self, fn, seqs, trace, *extra):
        cnt = len(seqs)
        blob = "".join(seqs).encode()
        offs = (ctypes.c_int64 * (cnt + 1))()
        toffs = (ctypes.c_int64 * (cnt + 1))()
        p = t = 0
        for q, s in enumerate(seqs):
            offs[q] = p
            toffs[q] = t
            p += len(s)
            t += 4 * len(s) + 8
        offs[cnt] = p
        toffs[cnt] = t
        res = (ctypes.c_int32 * max(cnt, 1))()
        tr = (Trace * max(t, 1))() if trace else None
        tl = (ctypes.c_int32 * max(cnt, 1))()
        self._check(fn(self.h, blob, offs, cnt, *extra, res, tr, toffs if trace else None, tl if trace else None))
        out = list(res[:cnt])
        if not trace:
            return out, None
        traces = [_tuples(tr[toffs[q]:toffs[q + 1]], tl[q]) for q in range(cnt)]
        return out, traces

    def zuker_batch(self, seqs, trace=True):
        return self._batch(self._L.dpx_zuker_batch, seqs, trace)

    def nussinov_batch(self, seqs, min_loop=0, trace=True):
        return self._batch(self._L.dpx_nussinov_batch, seqs, trace, ctypes.c_int(min_loop))

    def last_timing(self):
        f, z = ctypes.c_float(), ctypes.c_float()
        self._check(self._L.dpx_last_timing(self.h, ctypes.byref(f), ctypes.byref(z)))
        return f.value, z.value
