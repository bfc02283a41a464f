"""Python view of libdpx.so, the MI355X DP backend (C ABI in include/dpx.h).

The library is the product; this module only marshals arguments.  It fails
loudly when libdpx.so is missing: there is no CPU fallback.
"""
import ctypes
import os
import time

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "libdpx.so")
PAR_TURNER2004 = os.path.join(PKG, "params", "rna_turner2004.par")
# launch shapes chosen by the autotuner on MI355X (dpx_autotune), per kind and length bucket
TUNING_MI355X = os.path.join(PKG, "tuning", "mi355x.json")
# every parameter file shipped (ViennaRNA 2.0 format), by name
PAR_FILES = {"turner2004": PAR_TURNER2004,
             "turner1999": os.path.join(PKG, "params", "rna_turner1999.par"),
             "andronescu2007": os.path.join(PKG, "params", "rna_andronescu2007.par"),
             "dna_mathews2004": os.path.join(PKG, "params", "dna_mathews2004.par")}

OK, EINVAL, ENOMEM, EHIP, ENOPARAMS, EPARAMS, ETRACE, EBUSY, EAGAIN = range(9)

# exported symbols, as declared in include/dpx.h
EXPORTS = [
    "dpx_create", "dpx_destroy", "dpx_last_error", "dpx_set_rna_params",
    "dpx_nussinov", "dpx_zuker", "dpx_zuker_noalloc", "dpx_rnafold", "dpx_swat", "dpx_matchain",
    "dpx_zuker_batch", "dpx_nussinov_batch", "dpx_rnafold_batch", "dpx_last_timing", "dpx_now",
    "dpx_server_create", "dpx_server_destroy", "dpx_server_submit", "dpx_server_poll",
    "dpx_server_trace", "dpx_server_get_stats",
    "dpx_plan_create", "dpx_plan_run", "dpx_plan_timing", "dpx_plan_results", "dpx_plan_destroy",
    "dpx_tune_bucket", "dpx_get_tuning", "dpx_set_tuning", "dpx_autotune", "dpx_set_swat_trace_budget",
    "dpx_grammar_create", "dpx_grammar_destroy", "dpx_grammar_check", "dpx_grammar_get_info",
    "dpx_grammar_source", "dpx_grammar_parse", "dpx_grammar_parse_batch",
]

REQ_NUSSINOV, REQ_ZUKER, REQ_SWAT, REQ_MATCHAIN, REQ_RNAFOLD = 1, 2, 3, 4, 5
REQ_OF_KIND = {"nussinov": REQ_NUSSINOV, "zuker": REQ_ZUKER, "rnafold": REQ_RNAFOLD}
SUBMIT_TRACE, SUBMIT_TRACE_FROM_BEST = 1, 2


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


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_int64), ("slots", ctypes.c_int32), ("max_batch", ctypes.c_int32),
                ("max_pending", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("a", ctypes.c_char_p), ("n1", ctypes.c_int32),
                ("b", ctypes.c_char_p), ("n2", ctypes.c_int32),
                ("rows", ctypes.POINTER(ctypes.c_int32)), ("cols", ctypes.POINTER(ctypes.c_int32)),
                ("sc", SwatScores), ("min_loop", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("kind", ctypes.c_int32), ("status", ctypes.c_int32),
                ("score", ctypes.c_int32), ("best", ctypes.c_int32), ("best_i", ctypes.c_int32),
                ("best_j", ctypes.c_int32), ("trace_len", ctypes.c_int32), ("batch_size", ctypes.c_int32),
                ("cells", ctypes.c_int64), ("footprint", ctypes.c_int64),
                ("t_submit", ctypes.c_double), ("t_launch", ctypes.c_double), ("t_done", ctypes.c_double),
                ("device_ms", ctypes.c_float), ("reserved", ctypes.c_int32)]


TUNE_BUCKETS = 8
KIND_NAMES = {REQ_NUSSINOV: "nussinov", REQ_ZUKER: "zuker", REQ_SWAT: "swat", REQ_MATCHAIN: "matchain",
              REQ_RNAFOLD: "rnafold"}


class Tuning(ctypes.Structure):
    _fields_ = [("wg_per_cu", ctypes.c_int32), ("chunk", ctypes.c_int32), ("reserved", ctypes.c_int32 * 2)]


class TuneReport(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("bucket", ctypes.c_int32), ("n_lo", ctypes.c_int32),
                ("n_hi", ctypes.c_int32), ("count", ctypes.c_int32), ("t", Tuning), ("ms", ctypes.c_float),
                ("chosen", ctypes.c_int32)]


class GrammarInfo(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int32) for k in ("tracks", "n_tabs", "n_rules", "max_pos", "axiom_id", "window",
                                              "answer_bytes", "alphabet_bytes", "loops", "rna")] + \
               [("code_bytes", ctypes.c_int64)]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in
                ("submitted", "completed", "failed", "rejected", "launches", "cells_done",
                 "bytes_in_flight", "peak_bytes_in_flight", "max_batch_seen", "pending")]


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
        L.dpx_now.restype = ctypes.c_double
        L.dpx_server_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServerConfig), ctypes.POINTER(ctypes.c_void_p)]
        L.dpx_server_destroy.argtypes = [ctypes.c_void_p]
        L.dpx_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
        L.dpx_server_poll.argtypes = [ctypes.c_void_p, ctypes.POINTER(Result), ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int)]
        L.dpx_server_trace.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(Trace), ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int)]
        L.dpx_server_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServerStats)]
        L.dpx_plan_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64),
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.dpx_plan_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
        L.dpx_plan_results.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(Trace),
                                       ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32)]
        L.dpx_plan_destroy.argtypes = [ctypes.c_void_p]
        L.dpx_grammar_create.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]
        L.dpx_grammar_destroy.argtypes = [ctypes.c_void_p]
        L.dpx_grammar_check.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p)]
        L.dpx_grammar_get_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(GrammarInfo)]
        L.dpx_grammar_source.restype = ctypes.c_char_p
        L.dpx_grammar_source.argtypes = [ctypes.c_void_p]
        L.dpx_grammar_parse.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                        ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(Trace),
                                        ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        P64 = ctypes.POINTER(ctypes.c_int64)
        L.dpx_grammar_parse_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, P64, ctypes.c_char_p, P64,
                                              ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                                              ctypes.POINTER(Trace), P64, ctypes.POINTER(ctypes.c_int32)]
        _lib = L
    return _lib


def _tuples(buf, n):
    return [(t.i, t.j, t.rule, t.pos[0], t.pos[1]) for t in buf[:n]]


class Context:
    """One device context: resident energy tables and DP scratch."""

    def __init__(self, device=0, rna_params=PAR_TURNER2004, tuning=TUNING_MI355X):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.dpx_create(ctypes.byref(h), device)
        if rc != OK:
            raise DpxError(rc, "dpx_create(%d) failed" % device)
        self.h = h
        if rna_params:
            self.set_rna_params(rna_params)
        if tuning:
            self.load_tuning_file(tuning)

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

    def set_swat_trace_budget(self, nbytes):
        """Above this many bytes of traceback codes, swat() keeps checkpoint
        rows and recomputes segments during the traceback (same trace)."""
        self._check(self._L.dpx_set_swat_trace_budget(self.h, ctypes.c_int64(nbytes)))

    # ---- generic grammars (dpx.adp) ---------------------------------------
    def grammar(self, g):
        """Compiles a dpx.adp.Grammar for this device (hipRTC)."""
        return CompiledGrammar(self, g)

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

    def zuker_noalloc(self, seq, trace=True):
        """dpx_zuker_noalloc: dpx_zuker on the device's static scratch (n <= 1024)."""
        n = len(seq)
        cap = 4 * n + 8
        tr = (Trace * cap)() if trace else None
        score, tl = ctypes.c_int32(), ctypes.c_int()
        self._check(self._L.dpx_zuker_noalloc(self.h, seq.encode(), n, ctypes.byref(score), tr, cap,
                                              ctypes.byref(tl)))
        return score.value, (_tuples(tr, tl.value) if trace else None)

    def rnafold(self, seq, trace=True):
        n = len(seq)
        cap = 4 * n + 8
        tr = (Trace * cap)() if trace else None
        score, tl = ctypes.c_int32(), ctypes.c_int()
        self._check(self._L.dpx_rnafold(self.h, seq.encode(), n, ctypes.byref(score), tr, cap, ctypes.byref(tl)))
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
        t0 = time.perf_counter()
        rc = fn(self.h, blob, offs, cnt, *extra, res, tr, toffs if trace else None, tl if trace else None)
        self.last_call_s = time.perf_counter() - t0  # the C call alone: copies, kernels, staging
        self._check(rc)
        out = list(res[:cnt])
        if not trace:
            return out, None
        traces = [_tuples(tr[toffs[q]:toffs[q + 1]], tl[q]) for q in range(cnt)]
        return out, traces

    def zuker_batch(self, seqs, trace=True):
        return self._batch(self._L.dpx_zuker_batch, seqs, trace)

    def nussinov_batch(self, seqs, min_loop=0, trace=True):
        return self._batch(self._L.dpx_nussinov_batch, seqs, trace, ctypes.c_int(min_loop))

    def rnafold_batch(self, seqs, trace=True):
        return self._batch(self._L.dpx_rnafold_batch, seqs, trace)

    def plan(self, seqs, kind="zuker", min_loop=0, trace=True):
        """A batch uploaded once and replayed as a HIP graph (dpx_plan_*)."""
        return Plan(self, seqs, kind, min_loop, trace)

    def last_timing(self):
        f, z = ctypes.c_float(), ctypes.c_float()
        self._check(self._L.dpx_last_timing(self.h, ctypes.byref(f), ctypes.byref(z)))
        return f.value, z.value

    # ---- launch shape (dpx_*_tuning, dpx_autotune) -------------------------
    @staticmethod
    def _kind(kind):
        return kind if isinstance(kind, int) else {v: k for k, v in KIND_NAMES.items()}[kind]

    def get_tuning(self, kind, bucket):
        t = Tuning()
        self._check(self._L.dpx_get_tuning(self.h, self._kind(kind), bucket, ctypes.byref(t)))
        return {"wg_per_cu": t.wg_per_cu, "chunk": t.chunk}

    def set_tuning(self, kind, bucket, wg_per_cu=0, chunk=0):
        t = Tuning(wg_per_cu, chunk)
        self._check(self._L.dpx_set_tuning(self.h, self._kind(kind), bucket, ctypes.byref(t)))

    def tuning_table(self):
        """{kind name: [per-bucket tuning]} -- what save/load exchange as JSON."""
        return {name: [self.get_tuning(k, b) for b in range(TUNE_BUCKETS)] for k, name in KIND_NAMES.items()}

    def load_tuning_file(self, path):
        """A tuning table saved as JSON ({"table": {kind: [per-bucket shape]}})."""
        import json
        with open(path) as f:
            self.load_tuning_table(json.load(f)["table"])

    def load_tuning_table(self, table):
        for name, rows in table.items():
            for b, t in enumerate(rows):
                self.set_tuning(name, b, t["wg_per_cu"], t["chunk"])

    def autotune(self, kind, buckets=range(TUNE_BUCKETS), reps=2):
        """Time every launch shape on synthetic launches of each bucket and
        keep the fastest; returns one dict per candidate timed."""
        lo, hi = min(buckets), max(buckets)
        cap = 16 * (hi - lo + 1)
        rep = (TuneReport * cap)()
        n = ctypes.c_int()
        self._check(self._L.dpx_autotune(self.h, self._kind(kind), lo, hi, reps, rep, cap, ctypes.byref(n)))
        return [{"kind": KIND_NAMES[r.kind], "bucket": r.bucket, "n_lo": r.n_lo, "n_hi": r.n_hi, "count": r.count,
                 "wg_per_cu": r.t.wg_per_cu, "chunk": r.t.chunk, "ms": round(r.ms, 4), "chosen": bool(r.chosen)}
                for r in rep[:min(n.value, cap)]]


def grammar_check(spec_json):
    """(ok, generated HIP source, message) without a device."""
    src, msg = ctypes.c_char_p(), ctypes.c_char_p()
    rc = lib().dpx_grammar_check(spec_json.encode(), ctypes.byref(src), ctypes.byref(msg))
    return rc == OK, (src.value or b"").decode(), (msg.value or b"").decode()


class CompiledGrammar:
    """A grammar compiled and loaded on a context's device."""

    def __init__(self, ctx, g):
        self.ctx, self.g = ctx, g
        self._L = ctx._L
        h = ctypes.c_void_p()
        ctx._check(self._L.dpx_grammar_create(ctx.h, g.spec_json().encode(), ctypes.byref(h)))
        self.h = h
        self.info = GrammarInfo()
        self._L.dpx_grammar_get_info(self.h, ctypes.byref(self.info))

    @property
    def source(self):
        return self._L.dpx_grammar_source(self.h).decode()

    def parse(self, in1, in2=None, trace=True, cap=None):
        """(answer, found, trace records) of one input (two for two-track grammars)."""
        g = self.g
        b1 = g.pack(in1)
        b2 = g.pack(in2) if in2 is not None else b""
        n1 = len(b1) // self.info.alphabet_bytes
        n2 = len(b2) // self.info.alphabet_bytes
        cap = cap or 8 * (n1 + n2) + 64
        while True:
            tr = (Trace * cap)() if trace else None
            ans = ctypes.create_string_buffer(max(8, self.info.answer_bytes))
            found, tl = ctypes.c_int(), ctypes.c_int()
            rc = self._L.dpx_grammar_parse(self.ctx.h, self.h, b1, n1, b2 if in2 is not None else None, n2, ans,
                                           ctypes.byref(found), tr, cap if trace else 0, ctypes.byref(tl))
            if rc == ETRACE and trace and cap < (1 << 26):
                cap *= 4
                continue
            self.ctx._check(rc)
            a = g.unpack_answer(ans.raw)
            return a, bool(found.value), (_tuples(tr, tl.value) if trace else None)

    def parse_batch(self, inputs, trace=True, cap_per=None):
        """[(answer, found, trace)] for a list of inputs (pairs for two-track
        grammars), all instances in the same launches: one tile-queue launch
        (one track), one band-pipeline launch (two tracks without split
        loops) or one launch per anti-diagonal (two tracks with split loops)."""
        g = self.g
        two = g.tracks == 2
        b1 = [g.pack(x[0] if two else x) for x in inputs]
        b2 = [g.pack(x[1]) for x in inputs] if two else None
        ab = self.info.alphabet_bytes
        cnt = len(inputs)
        o1 = (ctypes.c_int64 * (cnt + 1))()
        o2 = (ctypes.c_int64 * (cnt + 1))()
        to = (ctypes.c_int64 * (cnt + 1))()
        p = p2 = t = 0
        for q in range(cnt):
            o1[q], o2[q], to[q] = p, p2, t
            p += len(b1[q]) // ab
            p2 += len(b2[q]) // ab if two else 0
            t += cap_per or 8 * (len(b1[q]) // ab + (len(b2[q]) // ab if two else 0)) + 64
        o1[cnt], o2[cnt], to[cnt] = p, p2, t
        ans = ctypes.create_string_buffer(max(8, self.info.answer_bytes * cnt))
        found = (ctypes.c_int32 * max(cnt, 1))()
        lens = (ctypes.c_int32 * max(cnt, 1))()
        tr = (Trace * max(t, 1))() if trace else None
        self.ctx._check(self._L.dpx_grammar_parse_batch(self.ctx.h, self.h, b"".join(b1), o1,
                                                        b"".join(b2) if two else None, o2 if two else None, cnt, ans,
                                                        found, tr, to if trace else None, lens if trace else None))
        out = []
        n = self.info.answer_bytes
        for q in range(cnt):
            a = g.unpack_answer(ans.raw[q * n:(q + 1) * n])
            out.append((a, bool(found[q]), _tuples(tr[to[q]:to[q + 1]], lens[q]) if trace else None))
        return out

    def close(self):
        if self.h:
            self._L.dpx_grammar_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def now():
    return lib().dpx_now()


class Server:
    """Continuous-batching request loop on one context (dpx_server_* in include/dpx.h)."""

    def __init__(self, ctx, hbm_budget=8 << 30, slots=3, max_batch=256, max_pending=1 << 20):
        self._L = ctx._L
        self.ctx = ctx
        cfg = ServerConfig(hbm_budget, slots, max_batch, max_pending, 0)
        h = ctypes.c_void_p()
        rc = self._L.dpx_server_create(ctx.h, ctypes.byref(cfg), ctypes.byref(h))
        if rc != OK:
            raise DpxError(rc, self._L.dpx_last_error(ctx.h).decode())
        self.h = h

    def close(self):
        if self.h:
            self._L.dpx_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _submit(self, req):
        rid = ctypes.c_uint64()
        rc = self._L.dpx_server_submit(self.h, ctypes.byref(req), ctypes.byref(rid))
        if rc != OK:
            raise DpxError(rc, "submit refused")
        return rid.value

    def submit_zuker(self, seq, trace=True):
        return self._submit(Request(kind=REQ_ZUKER, flags=SUBMIT_TRACE if trace else 0, a=seq.encode(), n1=len(seq)))

    def submit_rnafold(self, seq, trace=True):
        return self._submit(Request(kind=REQ_RNAFOLD, flags=SUBMIT_TRACE if trace else 0, a=seq.encode(),
                                    n1=len(seq)))

    def submit_nussinov(self, seq, min_loop=0, trace=True):
        return self._submit(Request(kind=REQ_NUSSINOV, flags=SUBMIT_TRACE if trace else 0, a=seq.encode(),
                                    n1=len(seq), min_loop=min_loop))

    def submit_matchain(self, rows, cols, trace=True):
        n = len(rows)
        r = (ctypes.c_int32 * n)(*rows)
        c = (ctypes.c_int32 * n)(*cols)
        return self._submit(Request(kind=REQ_MATCHAIN, flags=SUBMIT_TRACE if trace else 0, n1=n,
                                    rows=ctypes.cast(r, ctypes.POINTER(ctypes.c_int32)),
                                    cols=ctypes.cast(c, ctypes.POINTER(ctypes.c_int32))))

    def submit_swat(self, a, b, match, mismatch, gap_open, gap_ext, trace_from=0, trace=True):
        flags = (SUBMIT_TRACE if trace else 0) | (SUBMIT_TRACE_FROM_BEST if trace_from else 0)
        return self._submit(Request(kind=REQ_SWAT, flags=flags, a=a.encode(), n1=len(a), b=b.encode(), n2=len(b),
                                    sc=SwatScores(match, mismatch, gap_open, gap_ext)))

    def poll(self, max_results=256, wait_us=0):
        out = (Result * max_results)()
        got = ctypes.c_int()
        rc = self._L.dpx_server_poll(self.h, out, max_results, wait_us, ctypes.byref(got))
        if rc not in (OK, EAGAIN):
            raise DpxError(rc, "poll failed")
        return [out[q] for q in range(got.value)]

    def trace(self, rid, length):
        buf = (Trace * max(length, 1))()
        tl = ctypes.c_int()
        rc = self._L.dpx_server_trace(self.h, rid, buf, max(length, 1), ctypes.byref(tl))
        if rc == EAGAIN:
            return []
        if rc != OK:
            raise DpxError(rc, "trace fetch failed")
        return _tuples(buf, tl.value)

    def drop_trace(self, rid):
        tl = ctypes.c_int()
        self._L.dpx_server_trace(self.h, rid, None, 0, ctypes.byref(tl))

    def stats(self):
        st = ServerStats()
        self._L.dpx_server_get_stats(self.h, ctypes.byref(st))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def drain(self, ids, timeout_s=120.0, keep_trace=None):
        """Wait for every id in `ids`; returns {id: (Result, trace)}.  Traces
        are fetched for the ids in keep_trace (all when None), dropped otherwise."""
        want = set(ids)
        got = {}
        t_end = now() + timeout_s
        while want:
            if now() > t_end:
                raise DpxError(EAGAIN, "server did not finish %d requests in %.0f s" % (len(want), timeout_s))
            for r in self.poll(wait_us=20000):
                tr = []
                if r.trace_len > 0:
                    if keep_trace is None or r.id in keep_trace:
                        tr = self.trace(r.id, r.trace_len)
                    else:
                        self.drop_trace(r.id)
                got[r.id] = (r, tr)
                want.discard(r.id)
        return got


class Plan:
    """dpx_plan_*: a Zuker, Nussinov or RNAfold batch captured once as a HIP graph."""

    def __init__(self, ctx, seqs, kind="zuker", min_loop=0, trace=True):
        self._L = ctx._L
        self.ctx = ctx
        self.count = len(seqs)
        self.trace = trace
        blob = "".join(seqs).encode()
        offs = (ctypes.c_int64 * (self.count + 1))()
        self.toffs = (ctypes.c_int64 * (self.count + 1))()
        p = t = 0
        for q, s in enumerate(seqs):
            offs[q] = p
            self.toffs[q] = t
            p += len(s)
            t += 4 * len(s) + 8
        offs[self.count] = p
        self.toffs[self.count] = t
        self.total = t
        h = ctypes.c_void_p()
        k = REQ_OF_KIND[kind]
        ctx._check(self._L.dpx_plan_create(ctx.h, k, blob, offs, self.count, min_loop, 1 if trace else 0,
                                           ctypes.byref(h)))
        self.h = h

    def run(self):
        """Replay the graph; returns the device time in ms."""
        ms = ctypes.c_float()
        self.ctx._check(self._L.dpx_plan_run(self.h, ctypes.byref(ms)))
        return ms.value

    def timing(self):
        """(fill ms, finalize ms) of the last replay, from event nodes in the graph."""
        f, z = ctypes.c_float(), ctypes.c_float()
        self.ctx._check(self._L.dpx_plan_timing(self.h, ctypes.byref(f), ctypes.byref(z)))
        return f.value, z.value

    def results(self):
        res = (ctypes.c_int32 * self.count)()
        tr = (Trace * max(self.total, 1))() if self.trace else None
        tl = (ctypes.c_int32 * self.count)()
        self.ctx._check(self._L.dpx_plan_results(self.h, res, tr, self.toffs if self.trace else None,
                                                 tl if self.trace else None))
        out = list(res)
        if not self.trace:
            return out, None
        return out, [_tuples(tr[self.toffs[q]:self.toffs[q + 1]], tl[q]) for q in range(self.count)]

    def close(self):
        if self.h:
            self._L.dpx_plan_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
