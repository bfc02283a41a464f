"""Seeded synthetic inputs for the DP workloads.

A self-contained splitmix64 stream, so that every sequence set is the same
on any machine and numpy version.  Generators:
  rna_uniform(n, seed)        i.i.d. uniform over ACGU
  rna_gc(n, seed, gc=0.5)     i.i.d. with P(G)=P(C)=gc/2, P(A)=P(U)=(1-gc)/2
  dna_uniform(n, seed)        i.i.d. uniform over ACGT
  log_uniform_lengths(k, lo, hi, seed)
  matrix_chain(n, seed, lo=1, hi=1000)  consistent (rows, cols) pairs
"""
import math

import numpy as np

MASK = (1 << 64) - 1
GAMMA = 0x9E3779B97F4A7C15


def _mix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


class SplitMix64:
    def __init__(self, seed):
        # the seed is hashed, so nearby seeds start unrelated streams
        self.state = _mix((seed + 0x632BE59BD9B4E019) & MASK)

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        return _mix(self.state)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def below(self, m):
        return self.next() % m


def _stream(seed, n):
    """The first n outputs of SplitMix64(seed).next(), vectorised (uint64
    arithmetic wraps modulo 2**64 exactly like the scalar class)."""
    s0 = np.uint64(_mix((seed + 0x632BE59BD9B4E019) & MASK))
    with np.errstate(over="ignore"):
        z = s0 + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _letters(alphabet, idx):
    return np.frombuffer(alphabet.encode(), dtype=np.uint8)[idx].tobytes().decode()


def rna_uniform(n, seed):
    return _letters("ACGU", (_stream(seed, n) % np.uint64(4)).astype(np.int64))


def rna_gc(n, seed, gc=0.5):
    """P(G)=P(C)=gc/2, P(A)=P(U)=(1-gc)/2 from one uniform draw per base."""
    u = (_stream(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    idx = np.where(u < gc / 2, 0, np.where(u < gc, 1, np.where(u < gc + (1 - gc) / 2, 2, 3)))
    return _letters("GCAU", idx)


def dna_uniform(n, seed):
    return _letters("ACGT", (_stream(seed, n) % np.uint64(4)).astype(np.int64))


def dna_pair(n1, n2, seed):
    """Two sequences drawn back to back from one seeded stream."""
    s = dna_uniform(n1 + n2, seed)
    return s[:n1], s[n1:]


def log_uniform_lengths(k, lo, hi, seed):
    r = SplitMix64(seed)
    a, b = math.log(lo), math.log(hi)
    return [min(hi, max(lo, int(round(math.exp(a + (b - a) * r.uniform()))))) for _ in range(k)]


def rna_batch(k, lo, hi, seed, gc=0.5):
    """k sequences with log-uniform lengths in [lo,hi] (the batched-fold set)."""
    lens = log_uniform_lengths(k, lo, hi, seed)
    return [rna_gc(n, seed * 1000003 + q + 1, gc) for q, n in enumerate(lens)]


def matrix_chain(n, seed, lo=1, hi=1000):
    r = SplitMix64(seed)
    dims = [lo + r.below(hi - lo + 1) for _ in range(n + 1)]
    return dims[:-1], dims[1:]


# ---------------------------------------------------------------------------
# The reference's own input generators, restated so that its benchmark
# inputs can be reproduced exactly.

class JavaRandom:
    """java.util.Random (the generator behind scala.util.Random): 48-bit LCG,
    seed scrambled with 0x5DEECE66D; nextInt() = next(32) as signed int32."""

    MULT, ADD, MASK48 = 0x5DEECE66D, 0xB, (1 << 48) - 1

    def __init__(self, seed):
        self.seed = (seed ^ self.MULT) & self.MASK48

    def next(self, bits):
        self.seed = (self.seed * self.MULT + self.ADD) & self.MASK48
        r = self.seed >> (48 - bits)
        return r - (1 << bits) if r >= 1 << (bits - 1) else r

    def next_int(self):
        return self.next(32)


SCALA_UTILS_SEED = 123456748299  # Utils.reset (Utils.scala:7)


def _jabs_mod(x, m):
    """Math.abs(x) % m with JVM int semantics (abs(MIN_VALUE) stays negative;
    % truncates toward zero)."""
    a = abs(x) if x != -(1 << 31) else x
    r = abs(a) % m
    return -r if a < 0 else r


def scala_gen_rna(n, rnd):
    """Utils.genRNA (Utils.scala:29): n symbols of acgu from abs(nextInt) % 4."""
    return "".join("acgu"[_jabs_mod(rnd.next_int(), 4)] for _ in range(n))


def scala_gen_dna(n, rnd):
    """Utils.genDNA (Utils.scala:28): n symbols of ACGT."""
    return "".join("ACGT"[_jabs_mod(rnd.next_int(), 4)] for _ in range(n))


def scala_gen_mats(n, rnd):
    """Utils.genMats (Utils.scala:21-25): s = n draws, then one more draw d0;
    matrices (d0,s0), (s0,s1), ... ; each draw abs(nextInt) % 253 + 1.
    Returns (rows, cols)."""
    s = [_jabs_mod(rnd.next_int(), 253) + 1 for _ in range(n)]
    d0 = _jabs_mod(rnd.next_int(), 253) + 1
    rows = [d0] + s[:-1]
    return rows, s


class GlibcRandom:
    """glibc random()/srandom() with the default state (TYPE_3: additive
    feedback r[i] = r[i-3] + r[i-31] mod 2**32, output r >> 1; the state is
    seeded by r[i] = 16807 * r[i-1] mod (2**31 - 1) and 310 outputs are
    discarded), as used by the reference's bench2 harness."""

    def __init__(self, seed):
        seed &= 0xFFFFFFFF
        if seed == 0:
            seed = 1
        r = [0] * 34
        r[0] = seed - (1 << 32) if seed >= 1 << 31 else seed  # int32_t word
        for i in range(1, 31):
            # 16807 * r mod (2^31 - 1) by Schrage's method, as signed 32-bit words
            hi, lo = divmod(r[i - 1], 127773) if r[i - 1] >= 0 else (-((-r[i - 1]) // 127773), -((-r[i - 1]) % 127773))
            word = 16807 * lo - 2836 * hi
            if word < 0:
                word += 2147483647
            r[i] = word
        for i in range(31, 34):
            r[i] = r[i - 31]
        self.r = [x & 0xFFFFFFFF for x in r]
        for _ in range(310):
            self._step()

    def _step(self):
        v = (self.r[-31] + self.r[-3]) & 0xFFFFFFFF
        self.r.append(v)
        if len(self.r) > 64:
            del self.r[:-34]
        return v >> 1

    def random(self):
        return self._step()


BENCH2_SEED = 0x87654321  # srandom in bench2/wrapper.c:24 and gen.c:10


def bench2_seq(size, index, rnd=None):
    """Sequence number `index` (0-based) of the bench2 generator (gen.c:5-8,
    wrapper.c:14-21): `size` symbols "nacgu_"[random() % 4 + 1] each, after
    srandom(0x87654321); sequences are drawn back to back."""
    if rnd is None:
        rnd = GlibcRandom(BENCH2_SEED)
        for _ in range(index * size):
            rnd.random()
    return "".join("nacgu_"[rnd.random() % 4 + 1] for _ in range(size))


def served_stream(rate, seconds, seed=5, chain_share=0.1, zuker_share=0.8):
    """Seeded Poisson stream of the served mix: a list of (arrival s, kind,
    payload).  A matrix-chain request (64-512 matrices, dimensions uniform in
    [1,1000]) with probability chain_share; otherwise a Zuker fold (length
    log-uniform 30-3000, 50% GC) with probability zuker_share, else a Gotoh
    SW pair (each side log-uniform 500-20000 bp, ACGT)."""
    r = SplitMix64(seed)

    def log_uniform(lo, hi):
        return int(round(math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * r.uniform())))

    out, t, q = [], 0.0, 0
    while True:
        t += -math.log(1.0 - r.uniform()) / rate
        if t >= seconds:
            return out
        u = r.uniform()
        if u < chain_share:
            rows, cols = matrix_chain(64 + r.below(449), 500000 + q)
            out.append((t, "matchain", (rows, cols)))
        elif u < chain_share + (1 - chain_share) * zuker_share:
            out.append((t, "zuker", rna_gc(log_uniform(30, 3000), 600000 + q, 0.5)))
        else:
            a, b = dna_pair(log_uniform(500, 20000), log_uniform(500, 20000), 700000 + q)
            out.append((t, "swat", (a, b)))
        q += 1
