"""ctypes view of the CPU oracle (oracle/build/liboracle.so).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the parity reference or the timed CPU baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "build", "liboracle.so")
PAR_DIR = os.path.join(ROOT, "lamp-dp-mt_amd", "params")
PAR_PATH = os.path.join(PAR_DIR, "rna_turner2004.par")
# parameter files shipped beside Turner 2004 (the reference's src/librna/vienna/*.par)
PAR_SETS = {"turner2004": "rna_turner2004.par", "turner1999": "rna_turner1999.par",
            "andronescu2007": "rna_andronescu2007.par", "dna_mathews2004": "dna_mathews2004.par"}

CODE = {"A": 1, "C": 2, "G": 3, "U": 4, "T": 4}


class Trace(ctypes.Structure):
    _fields_ = [("i", ctypes.c_int32), ("j", ctypes.c_int32), ("rule", ctypes.c_int32),
                ("pos", ctypes.c_int32 * 2)]


_lib = None
_params = {}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "build/liboracle.so"])
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_params_size = None
    return _lib


def params(par=None):
    """Scaled parameters of one .par file (Turner 2004 by default, else a
    PAR_SETS name) as an opaque buffer."""
    name = par or "turner2004"
    if name not in _params:
        buf = ctypes.create_string_buffer(1 << 20)
        rc = lib().orc_load_params(os.path.join(PAR_DIR, PAR_SETS[name]).encode(), buf)
        if rc != 0:
            raise RuntimeError("orc_load_params failed: %d" % rc)
        _params[name] = buf
    return _params[name]


def encode(seq):
    return (ctypes.c_uint8 * len(seq))(*[CODE[c] for c in seq.upper()])


def _trace_list(tr, n, tl):
    out = []
    for q in range(n):
        t = tr[q]
        out.append((t.i, t.j, t.rule, t.pos[0], t.pos[1]))
    return out


def zuker(seq, trace=True, par=None):
    n = len(seq)
    cap = 4 * (n + 2)
    tr = (Trace * cap)()
    tl = ctypes.c_int(0)
    f = lib().orc_zuker
    f.restype = ctypes.c_int32
    score = f(params(par), encode(seq), n, tr if trace else None, cap, ctypes.byref(tl), None)
    return score, (_trace_list(tr, tl.value, tl) if trace else None)


def zuker_tables(seq):
    n = len(seq)
    tab = np.zeros((n + 1, n + 1, 4), dtype=np.int32)
    f = lib().orc_zuker
    f.restype = ctypes.c_int32
    score = f(params(), encode(seq), n, None, 0, None, tab.ctypes.data_as(ctypes.c_void_p))
    return score, tab


def nussinov(seq, min_loop=0, trace=True):
    n = len(seq)
    cap = 4 * (n + 2)
    tr = (Trace * cap)()
    tl = ctypes.c_int(0)
    f = lib().orc_nussinov
    f.restype = ctypes.c_int32
    score = f(encode(seq), n, min_loop, tr if trace else None, cap, ctypes.byref(tl), None)
    return score, (_trace_list(tr, tl.value, tl) if trace else None)


def swat(a, b, match, mismatch, gap_open, gap_ext, trace=True, from_best=False):
    n1, n2 = len(a), len(b)
    cap = 2 * (n1 + n2) + 4
    tr = (Trace * cap)()
    tl = ctypes.c_int(0)
    best = ctypes.c_int32(0)
    bi = ctypes.c_int(0)
    bj = ctypes.c_int(0)
    f = lib().orc_swat
    f.restype = ctypes.c_int32
    score = f(encode(a), n1, encode(b), n2, match, mismatch, gap_open, gap_ext,
              ctypes.byref(best), ctypes.byref(bi), ctypes.byref(bj), 1 if from_best else 0,
              tr if trace else None, cap, ctypes.byref(tl))
    return score, (best.value, bi.value, bj.value), (_trace_list(tr, tl.value, tl) if trace else None)


def matchain(rows, cols, trace=True):
    n = len(rows)
    cap = 2 * n + 4
    tr = (Trace * cap)()
    tl = ctypes.c_int(0)
    r = (ctypes.c_int32 * n)(*rows)
    c = (ctypes.c_int32 * n)(*cols)
    f = lib().orc_matchain
    f.restype = ctypes.c_int32
    score = f(r, c, n, tr if trace else None, cap, ctypes.byref(tl))
    return score, (_trace_list(tr, tl.value, tl) if trace else None)


RF_EMPTY = 999999


def rnafold(seq, trace=True, par=None):
    n = len(seq)
    cap = 8 * (n + 2)
    tr = (Trace * cap)()
    tl = ctypes.c_int(0)
    f = lib().orc_rnafold
    f.restype = ctypes.c_int32
    score = f(params(par), encode(seq), n, tr if trace else None, cap, ctypes.byref(tl), None)
    return score, (_trace_list(tr, tl.value, tl) if trace else None)


class Rna(ctypes.Structure):
    _fields_ = [("P", ctypes.c_void_p), ("s", ctypes.c_void_p), ("n", ctypes.c_int)]


def energy_fns(seq):
    """(hl_energy, il_energy, sr_energy) of the oracle bound to seq."""
    codes = encode(seq)
    r = Rna(ctypes.cast(params(), ctypes.c_void_p), ctypes.cast(codes, ctypes.c_void_p), len(seq))
    L = lib()
    keep = (codes, r)

    def hl(i, j):
        return L.orc_hl_energy(ctypes.byref(keep[1]), i, j)

    def il(i, k, l, j):
        return L.orc_il_energy(ctypes.byref(keep[1]), i, k, l, j)

    def sr(i, j):
        return L.orc_sr_energy(ctypes.byref(keep[1]), i, j)

    return hl, il, sr
