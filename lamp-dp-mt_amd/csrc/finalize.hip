// finalize.hip -- per-instance answer and backtrace for the triangular grammars.
//
// One workgroup per instance.  The fill kernels store values only; here the
// backtrace walks the reference's BFS (bench2/zu.c:49-93, nu.c:38-59,
// CodeGen.scala:233-252) and, for every visited cell, re-evaluates the
// candidates of its table in the reference's evaluation order, all 256
// threads in parallel, taking the first candidate equal to the stored
// optimum.  Because the reference keeps a candidate only on strict
// improvement, that first candidate is exactly its stored back-pointer.
// Zuker's exterior table st (only column j = n exists, zu.c:37 fun14) is a
// 1-D O(n^2) pass computed here before the backtrace.
#include <hip/hip_runtime.h>

#include "engine.h"
#include "rna_model.h"

namespace dpx {

typedef struct { int32_t i, j, rule, pos[2]; } TraceRec;

__device__ __forceinline__ int fmin_(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int fmax_(int a, int b) { return a > b ? a : b; }

constexpr unsigned long long fbp_word(int lo) {
  unsigned long long w = 0;
  for (int q = 0; q < 16; ++q) {
    int idx = lo + q, x = idx / 5, y = idx % 5, v = 7;
    if (x == 1 && y == 4) v = 5;
    if (x == 2 && y == 3) v = 1;
    if (x == 3 && y == 2) v = 2;
    if (x == 3 && y == 4) v = 3;
    if (x == 4 && y == 3) v = 4;
    if (x == 4 && y == 1) v = 6;
    if (idx < 25) w |= (unsigned long long)v << (4 * q);
  }
  return w;
}
__device__ __forceinline__ int fbpc(int x, int y) {
  const int idx = x * 5 + y;
  const unsigned long long w = idx < 16 ? fbp_word(0) : fbp_word(16);
  return (int)((w >> (4 * (idx & 15))) & 15);
}

struct Cells {
  const int32_t* tab;
  int nb, NTv;
  int64_t tile0;
  __device__ int get(int t, int i, int j) const {
    const int I = i / T, J = j / T;
    return tab[(size_t)(tile0 + tile_index(nb, I, J)) * (size_t)(NTv * T * T) + t * T * T + (i % T) * T + (j % T)];
  }
};

// BFS backtrace driver.  The reference takes its queued records one at a
// time (zu.c:49-93, CodeGen.scala:233-252); a record's rule depends only on
// its own cell, so each round here resolves up to NW queued records at once
// and then appends their children in record order: the same records, in
// the same order.  With GROUPS the NW waves are split into groups of
// g = NW / m (rounded down to a power of two) for the m records of the
// round, so a lone record is searched by all NW waves (worth it where a
// record has hundreds of costly candidates, as Zuker's cl, and the
// workgroup has 16 waves: launches of few instances); otherwise one wave
// searches one record.
// rule_of(rec, match, &rule, &p0, &p1) is called by every wave once per
// round and calls match(ncand, target, cand) exactly once: match returns
// the smallest candidate index c < ncand with cand(c, &v) true and
// v == target (-1 if none) for the wave's record; with GROUPS it contains a
// workgroup barrier.  rule_of returns false when no candidate matches.  kids(rec,
// push) appends the children of a resolved record.  *len receives the
// record count, or -1 (no match, or more records than cap).
template <int NW, bool GROUPS, class RuleOf, class Kids>
__device__ void bfs_rounds(TraceRec* tr, int cap, TraceRec root, RuleOf&& rule_of, Kids&& kids, int32_t* len) {
  constexpr int NONE = 0x7fffffff;
  __shared__ int s_rd, s_wr, s_over;
  __shared__ int s_fail[NW];  // per record: no candidate matched it this round
  __shared__ int s_hit[NW];   // per record: smallest matching candidate found so far
  __shared__ TraceRec s_rec[NW];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (threadIdx.x == 0) {
    s_rd = 0;
    s_wr = 0;
    s_over = 0;
    for (int q = 0; q < NW; ++q) s_hit[q] = NONE;
    if (cap > 0) {
      tr[0] = root;
      s_wr = 1;
    } else {
      s_over = 1;
    }
  }
  __syncthreads();
  for (;;) {
    const int rd = s_rd, wr = s_wr;
    if (rd >= wr || s_over) break;
    const int m = wr - rd < NW ? wr - rd : NW;
    int g = GROUPS ? NW : 1;
    while (g > 1 && g * m > NW) g >>= 1;
    const int rw = w / g, sub = w % g;
    const bool act = rw < m;
    // the g waves of a group take the 64-candidate chunks sub, sub + g, ...
    // of their record; a wave stops at its first hit or once a hit below
    // its next chunk is known; the smallest hit is the first match
    auto match = [&](int ncand, int target, auto&& cand) -> int {
      if (!GROUPS) {
        if (act)
          for (int base = 0; base < ncand; base += 64) {
            const int c = base + lane;
            int v = 0;
            const bool hit = c < ncand && cand(c, &v) && v == target;
            const unsigned long long mk = __ballot(hit);
            if (mk) return base + (int)__ffsll(mk) - 1;
          }
        return -1;
      }
      if (act) {
        const volatile int* hp = &s_hit[rw];
        for (int it = 0;; ++it) {
          const int base = (it * g + sub) * 64;
          if (base >= ncand || base > *hp) break;
          const int c = base + lane;
          int v = 0;
          const bool hit = c < ncand && cand(c, &v) && v == target;
          const unsigned long long mk = __ballot(hit);
          if (mk) {
            if (lane == 0) atomicMin(&s_hit[rw], base + (int)__ffsll(mk) - 1);
            break;
          }
        }
      }
      __syncthreads();
      const int h = act ? s_hit[rw] : NONE;
      return h == NONE ? -1 : h;
    };
    TraceRec cur = tr[rd + (act ? rw : 0)];  // waves without a record search nothing
    int rule = -1, p0 = 0, p1 = 0;
    const bool ok = rule_of(cur, match, &rule, &p0, &p1);
    if (act && sub == 0 && lane == 0) {
      // s_over is read by every wave at the top of the round, so a failure
      // is recorded per record and folded in below, between barriers
      s_fail[rw] = ok ? 0 : 1;
      cur.rule = rule;
      cur.pos[0] = p0;
      cur.pos[1] = p1;
      tr[rd + rw] = cur;
      s_rec[rw] = cur;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int q = 0; q < m; ++q)
        if (s_fail[q]) s_over = 1;
      for (int q = 0; q < NW; ++q) s_hit[q] = NONE;
      int wr2 = wr;
      auto push = [&](int pi, int pj, int pr) {
        if (wr2 >= cap) {
          s_over = 1;
          return;
        }
        tr[wr2++] = TraceRec{pi, pj, pr, {0, 0}};
      };
      for (int q = 0; q < m && !s_over; ++q) kids(s_rec[q], push);
      s_rd = rd + m;
      s_wr = wr2;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *len = s_over ? -1 : s_wr;
}

// ----------------------------------------------------------------- Nussinov
template <int NT>
__global__ __launch_bounds__(NT) void nu_finalize(TriArgs a) {
  const TriInst in = a.inst[blockIdx.x];
  const Cells C{a.tab, in.nb, NT_NU, in.tile0};
  const uint8_t* s = a.codes + in.seq_off;
  const int n = in.n;
  if (threadIdx.x == 0) a.results[in.out] = C.get(0, 0, n);
  if (!a.want_trace) return;
  auto rule_of = [&](const TraceRec& cur, auto&& match, int* rule, int* p0, int* p1) -> bool {
    const int i = cur.i, j = cur.j;
    const int target = C.get(0, i, j);
    // candidates in nu_kern.h order: nil, left, right, pair, split(k asc)
    auto cand = [&](int c, int* v) -> bool {
      if (c == 0) { *v = 0; return i == j; }
      if (c == 1) { if (i + 1 > j) return false; *v = C.get(0, i + 1, j); return true; }
      if (c == 2) { if (i + 1 > j) return false; *v = C.get(0, i, j - 1); return true; }
      if (c == 3) {
        if (i + 2 + a.min_loop > j || fbpc(s[i], s[j - 1]) == 7) return false;
        *v = C.get(0, i + 1, j - 1) + 1;
        return true;
      }
      const int k = i + 1 + (c - 4);
      if (k > j - 1) return false;
      *v = C.get(0, i, k) + C.get(0, k, j);
      return true;
    };
    const int c = match(4 + fmax_(0, j - i - 1), target, cand);
    *rule = c < 4 ? c : 4;
    *p0 = c >= 4 ? i + 1 + (c - 4) : 0;
    *p1 = 0;
    return c >= 0;
  };
  auto kids = [&](const TraceRec& e, auto&& push) {
    const int i = e.i, j = e.j;
    switch (e.rule) {
      case 1: push(i + 1, j, 0); break;
      case 2: push(i, fmax_(i, j - 1), 0); break;
      case 3: push(i + 1, fmax_(i + 1, j - 1), 0); break;
      case 4: push(i, e.pos[0], 0); push(e.pos[0], j, 0); break;
      default: break;
    }
  };
  bfs_rounds<NT / 64, false>((TraceRec*)a.traces + in.trace_off, in.trace_cap, TraceRec{0, n, 0, {0, 0}}, rule_of, kids,
             a.trace_lens + in.out);
}

// ------------------------------------------------------------- matrix chain
template <int NT>
__global__ __launch_bounds__(NT) void mc_finalize(TriArgs a) {
  const TriInst in = a.inst[blockIdx.x];
  const Cells C{a.tab, in.nb, NT_MC, in.tile0};
  const int32_t* rows = a.dims + in.seq_off;
  const int32_t* cols = rows + in.n;
  const int n = in.n;
  if (threadIdx.x == 0) a.results[in.out] = C.get(0, 0, n);
  if (!a.want_trace) return;
  auto rule_of = [&](const TraceRec& cur, auto&& match, int* rule, int* p0, int* p1) -> bool {
    const int i = cur.i, j = cur.j;
    const int target = C.get(0, i, j);
    auto cand = [&](int c, int* v) -> bool {
      if (c == 0) { *v = 0; return i + 1 == j; }
      const int k = i + c;
      if (k > j - 1) return false;
      *v = (int)((uint32_t)C.get(0, i, k) + (uint32_t)C.get(0, k, j) +
                 (uint32_t)rows[i] * (uint32_t)rows[k] * (uint32_t)cols[j - 1]);
      return true;
    };
    const int c = match(1 + fmax_(0, j - i - 1), target, cand);
    *rule = c == 0 ? 0 : 1;
    *p0 = c <= 0 ? 0 : i + c;
    *p1 = 0;
    return c >= 0;
  };
  auto kids = [&](const TraceRec& e, auto&& push) {
    if (e.rule == 1) {
      push(e.i, e.pos[0], 0);
      push(e.pos[0], e.j, 0);
    }
  };
  bfs_rounds<NT / 64, false>((TraceRec*)a.traces + in.trace_off, in.trace_cap, TraceRec{0, n, 0, {0, 0}}, rule_of, kids,
             a.trace_lens + in.out);
}

// -------------------------------------------------------------------- Zuker
// the energy tables' small head (every field before int21, 12 KB): the
// backtrace's candidate energies read it from LDS, staged once per
// workgroup, instead of through chains of dependent global loads
constexpr int kRnaHeadWords = (int)(offsetof(DevRna, int21) / sizeof(int32_t));
static_assert(offsetof(DevRna, int21) % sizeof(int32_t) == 0, "energy table head");
__device__ __forceinline__ const DevRna* stage_rna_head(const void* g, int32_t* lds) {
  for (int e = threadIdx.x; e < kRnaHeadWords; e += blockDim.x) lds[e] = ((const int32_t*)g)[e];
  __syncthreads();
  return (const DevRna*)lds;  // valid for the fields before int21
}

struct ZuF {
  const uint8_t* s;
  const DevRna* P;  // the head in LDS (fields before int21)
  const DevRna* Q;  // the whole tables in memory (int21, int22)
  const int32_t* hl;
  int n;
  __device__ int S(int k) const { return s[k]; }
  __device__ int bp(int x, int y) const { return fbpc(S(x), S(y)); }
  __device__ int termau(int x, int y) const {
    const int p = S(x), q = S(y);
    return ((p == 3 && q == 2) || (p == 2 && q == 3)) ? 0 : P->term_au;
  }
  __device__ int sr(int i, int j) const { return P->stack[bp(i, j)][bp(j - 1, i + 1)]; }
  __device__ int hairpin(int i, int j) const {
    const int size = j - i - 1;
    if (size < 3) return 600;
    const int cls = size == 3 ? 0 : size == 4 ? 1 : size == 6 ? 2 : -1;
    if (cls >= 0) {
      int key = 0;
      for (int m = 0; m <= size + 1; ++m) key = key * 8 + S(i + m);
      for (int e = 0; e < P->n_loop[cls]; ++e)
        if (P->loop_key[cls][e] == key) return P->loop_e[cls][e];
    }
    const int ent = hl[size];
    if (size == 3) return ent + termau(i, j);
    return ent + P->mmH[bp(i, j)][S(i + 1)][S(j - 1)];
  }
  __device__ int il(int i, int k, int l, int j) const {
    const int sl = k - i - 1, sr_ = j - l - 1;
    const int ocb = bp(i, j), icb = bp(l, k);
    if (sl == 1) {
      if (sr_ == 1) return P->int11[ocb][bp(j - 2, i + 2)][S(i + 1)][S(j - 1)];
      if (sr_ == 2) return Q->int21[ocb][bp(j - 3, i + 2)][S(i + 1)][S(j - 2)][S(j - 1)];
      return P->F[sl][sr_] + P->mm1nI[ocb][S(i + 1)][S(j - 1)] + P->mm1nI[icb][S(l + 1)][S(k - 1)];
    }
    if (sl == 2) {
      if (sr_ == 1) return Q->int21[bp(j - 2, i + 3)][ocb][S(j - 1)][S(i + 1)][S(i + 2)];
      if (sr_ == 2) return Q->int22[ocb][bp(j - 3, i + 3)][S(i + 1)][S(i + 2)][S(j - 2)][S(j - 1)];
      if (sr_ == 3) return P->il5 + P->ninio0 + P->mm23I[ocb][S(i + 1)][S(j - 1)] + P->mm23I[icb][S(l + 1)][S(k - 1)];
    } else if (sl == 3 && sr_ == 2) {
      return P->il5 + P->ninio0 + P->mm23I[ocb][S(i + 1)][S(j - 1)] + P->mm23I[icb][S(l + 1)][S(k - 1)];
    } else if (sr_ == 1) {
      return P->F[sl][sr_] + P->mm1nI[ocb][S(i + 1)][S(j - 1)] + P->mm1nI[icb][S(l + 1)][S(k - 1)];
    }
    return P->F[sl][sr_] + P->mmI[ocb][S(i + 1)][S(j - 1)] + P->mmI[icb][S(l + 1)][S(k - 1)];
  }
  __device__ int bulge_l(int i, int j, int k) const {
    const int size = k - i - 2;
    if (size == 1) return P->bulge[1] + P->stack[bp(i + 1, j - 2)][bp(j - 3, k)];
    return P->bulge[size] + termau(i + 1, j - 2) + termau(j - 3, k);
  }
  __device__ int bulge_r(int i, int j, int k) const {
    const int size = j - k - 2;
    if (size == 1) return P->bulge[1] + P->stack[bp(i + 1, j - 2)][bp(k - 1, i + 2)];
    return P->bulge[size] + termau(i + 1, j - 2) + termau(k - 1, i + 2);
  }
};

template <int NT>
__global__ __launch_bounds__(NT) void zu_finalize(TriArgs a) {
  extern __shared__ int st[];  // st(i,n), i = 0..n
  const TriInst in = a.inst[blockIdx.x];
  const Cells C{a.tab, in.nb, NT_ZU, in.tile0};
  const int n = in.n;
  __shared__ int32_t s_head[kRnaHeadWords];
  uint8_t* s_seq = (uint8_t*)(st + n + 1);  // the sequence, after the st column
  for (int e = threadIdx.x; e < n; e += NT) s_seq[e] = a.codes[in.seq_off + e];
  const ZuF z{s_seq, stage_rna_head(a.rna, s_head), (const DevRna*)a.rna, a.hl_ent, n};
  const int mlint = z.P->ml_intern;
  // ---- st column (zu_kern.h:166-182): rule 12 st(i+1), rule 13 D(i,k)+st(k), rule 14 i==n.
  // Rows are taken in blocks of 32 from the bottom.  For a block [lo, hi] the
  // split points k > hi (st(k) already known) are reduced by all threads at
  // once; the split points inside the block are then walked row by row by
  // one wave from D values staged in LDS.
  __shared__ int far[T], nd[T][T + 1];
  if (threadIdx.x == 0) st[n] = 0;
  __syncthreads();  // st(n) is read by every thread of the first block's far pass
  const int tid = threadIdx.x, lane = tid & 31;
  for (int hi = n - 1; hi >= 0; hi -= T) {
    const int lo = fmax_(0, hi - T + 1);
    for (int pass = 0; pass < T * 32 / NT; ++pass) {  // NT / 32 rows at a time, 32 lanes each
      const int r = pass * (NT / 32) + (tid >> 5), i = lo + r;
      int m = INV;
      if (i <= hi) {
        // sixteen split points per lane in flight at a time
        const int k0 = fmax_(hi + 1, i + 7) + lane;
        for (int kb = k0; kb <= n; kb += 16 * 32) {
          int u[16];
#pragma unroll
          for (int x = 0; x < 16; ++x) {
            const int k = kb + 32 * x;
            u[x] = k <= n ? C.get(ZT_U, i, k) : INV;
          }
#pragma unroll
          for (int x = 0; x < 16; ++x) {
            const int k = kb + 32 * x;
            if (u[x] < INV_THR) m = fmin_(m, u[x] - mlint + st[k]);
          }
        }
      }
      m = dpp_half_min(m);
      if (lane == 31) far[r] = m;
    }
    {
      // every global read of the block before its LDS writes (one latency)
      constexpr int NQ = (T * T + NT - 1) / NT;
      int v[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int e = tid + q * NT, r = e / T, c = e % T, i = lo + r, k = lo + c;
        v[q] = INV;
        if (e < T * T && i <= hi && k <= hi && k >= i + 7) {
          const int U = C.get(ZT_U, i, k);
          if (U < INV_THR) v[q] = U - mlint;
        }
      }
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (tid + q * NT < T * T) nd[(tid + q * NT) / T][(tid + q * NT) % T] = v[q];
    }
    __syncthreads();
    if (tid < 64) {
      for (int i = hi; i >= lo; --i) {
        const int r = i - lo, k = lo + lane;
        int m = INV;
        if (tid < 32 && k >= i + 7 && k <= hi && nd[r][lane] < INV_THR) m = nd[r][lane] + st[k];
        m = dpp_half_min(m);  // on lanes 16..31
        if (tid == 31) st[i] = fmin_(fmin_(st[i + 1], far[r]), m);
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): st[i] visible to the wave's next row
        __builtin_amdgcn_wave_barrier();
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) a.results[in.out] = st[0];
  if (!a.want_trace) return;
  auto rule_of = [&](const TraceRec& cur, auto&& match, int* rule, int* p0, int* p1) -> bool {
    const int i = cur.i, j = cur.j, tabrule = cur.rule, d = j - i;
    int ncand, target;
    if (tabrule == 12) { target = st[i]; ncand = 2 + fmax_(0, n - (i + 7) + 1); }
    else if (tabrule == 0) { target = C.get(ZT_CL, i, j); ncand = 963; }
    else if (tabrule == 6) { target = C.get(ZT_ML, i, j); ncand = 1 + fmax_(0, j - 7 - (i + 7) + 1); }
    else { target = C.get(ZT_ML1, i, j); ncand = 2 + fmax_(0, j - 7 - (i + 7) + 1) + fmax_(0, j - 1 - (i + 7) + 1); }
    const int nk9 = fmax_(0, j - 7 - (i + 7) + 1);
    auto cand = [&](int c, int* v) -> bool {
      if (tabrule == 12) {  // st
        if (c == 0) { if (i + 1 > n) return false; *v = st[i + 1]; return true; }
        const int k = i + 7 + (c - 1);
        if (c - 1 < fmax_(0, n - (i + 7) + 1)) {
          const int U = C.get(ZT_U, i, k);
          if (U >= INV_THR) return false;
          *v = U - mlint + st[k];
          return true;
        }
        *v = 0;
        return i == n;
      }
      if (tabrule == 0) {  // cl: rules 0..5
        const int sr = z.sr(i, j - 1);
        if (c == 0) {
          if (d < 9) return false;
          const int x = C.get(ZT_CL, i + 1, j - 1);
          if (x >= INV_THR) return false;
          *v = x + sr; return true;
        }
        if (c == 1) { if (d < 7) return false; *v = z.hairpin(i + 1, j - 2) + sr; return true; }
        if (c < 32) {
          const int k = i + 3 + (c - 2);
          if (d < 12 || k > fmin_(j - 9, i + 32)) return false;
          const int x = C.get(ZT_CL, k, j - 2);
          if (x >= INV_THR) return false;
          *v = x + z.bulge_l(i, j, k) + sr; return true;
        }
        if (c < 62) {
          const int k = j - 32 + (c - 32);
          if (d < 12 || k < i + 9) return false;
          const int x = C.get(ZT_CL, i + 2, k);
          if (x >= INV_THR) return false;
          *v = x + z.bulge_r(i, j, k) + sr; return true;
        }
        if (c < 962) {
          const int kk = (c - 62) / 30, ll = (c - 62) % 30;
          const int k = j - 32 + kk, l = i + 3 + ll;
          if (d < 13 || k < i + 10 || l > k - 7) return false;
          const int x = C.get(ZT_CL, l, k);
          if (x >= INV_THR) return false;
          *v = x + z.il(i + 1, l, k - 1, j - 2) + sr; return true;
        }
        if (d < 18) return false;
        const int x = C.get(ZT_ML, i + 2, j - 2);
        if (x >= INV_THR) return false;
        *v = z.P->ml_closing + mlint + x + z.termau(i + 1, j - 2) + sr +
             z.P->mmM[z.bp(j - 2, i + 1)][z.S(j - 3)][z.S(i + 2)];
        return true;
      }
      if (tabrule == 6) {  // ml: rules 6, 7
        if (c == 0) {
          if (d < 15) return false;
          const int x = C.get(ZT_ML, i + 1, j);
          if (x >= INV_THR) return false;
          *v = x; return true;
        }
        const int k = i + 7 + (c - 1);
        const int U = C.get(ZT_U, i, k), M1 = C.get(ZT_ML1, k, j);
        if (U >= INV_THR || M1 >= INV_THR) return false;
        *v = U + M1; return true;
      }
      // ml1: rules 8, 9, 10, 11
      if (c == 0) {
        if (d < 8) return false;
        const int x = C.get(ZT_ML1, i + 1, j);
        if (x >= INV_THR) return false;
        *v = x; return true;
      }
      if (c <= nk9) {
        const int k = i + 7 + (c - 1);
        const int U = C.get(ZT_U, i, k), M1 = C.get(ZT_ML1, k, j);
        if (U >= INV_THR || M1 >= INV_THR) return false;
        *v = U + M1; return true;
      }
      if (c == nk9 + 1) {
        if (d < 7) return false;
        const int U = C.get(ZT_U, i, j);
        if (U >= INV_THR) return false;
        *v = U; return true;
      }
      const int k = i + 7 + (c - nk9 - 2);
      if (k > j - 1) return false;
      const int U = C.get(ZT_U, i, k);
      if (U >= INV_THR) return false;
      *v = U; return true;
    };
    const int c = match(ncand, target, cand);
    if (c < 0) {
      // no candidate: the record is dropped with the whole trace
    } else if (tabrule == 12) {
      if (c == 0) *rule = 12;
      else if (c - 1 < fmax_(0, n - (i + 7) + 1)) { *rule = 13; *p0 = i + 7 + (c - 1); }
      else *rule = 14;
    } else if (tabrule == 0) {
      if (c == 0) *rule = 0;
      else if (c == 1) *rule = 1;
      else if (c < 32) { *rule = 2; *p0 = i + 3 + (c - 2); }
      else if (c < 62) { *rule = 3; *p0 = j - 32 + (c - 32); }
      else if (c < 962) { *rule = 4; *p0 = i + 3 + (c - 62) % 30; *p1 = j - 32 + (c - 62) / 30; }
      else *rule = 5;
    } else if (tabrule == 6) {
      if (c == 0) *rule = 6; else { *rule = 7; *p0 = i + 7 + (c - 1); }
    } else {
      if (c == 0) *rule = 8;
      else if (c <= nk9) { *rule = 9; *p0 = i + 7 + (c - 1); }
      else if (c == nk9 + 1) *rule = 10;
      else { *rule = 11; *p0 = i + 7 + (c - nk9 - 2); }
    }
    return c >= 0;
  };
  auto kids = [&](const TraceRec& e, auto&& push) {
    const int i = e.i, j = e.j;
    switch (e.rule) {  // children as zu.c:77-91
      case 0: push(i + 1, fmax_(i + 8, j - 1), 0); break;
      case 2: push(e.pos[0], fmax_(i + 10, fmax_(i + 11, j - 1) - 1), 0); break;
      case 3: push(i + 2, e.pos[0], 0); break;
      case 4: push(e.pos[0], e.pos[1], 0); break;
      case 5: push(i + 2, fmax_(i + 16, fmax_(i + 17, j - 1) - 1), 6); break;
      case 6: push(i + 1, j, 6); break;
      case 7: case 9: case 13:
        push(i, fmax_(i + 7, e.pos[0]), 0); push(e.pos[0], j, e.rule == 13 ? 12 : 8); break;
      case 8: push(i + 1, j, 8); break;
      case 10: push(i, fmax_(i + 7, j), 0); break;
      case 11: push(i, fmax_(i + 7, e.pos[0]), 0); break;
      case 12: push(i + 1, j, 12); break;
      default: break;
    }
  };
  bfs_rounds<NT / 64, (NT >= 1024)>((TraceRec*)a.traces + in.trace_off, in.trace_cap, TraceRec{0, n, 12, {0, 0}}, rule_of, kids,
             a.trace_lens + in.out);
}

// ------------------------------------------------------------------ RNAfold
// Interior loop of outer pair (i,j) and inner pair (k,l) for every size
// split, bulges and the stacked pair included (librna_impl.h:199-265).
__device__ int rf_il(const ZuF& z, int i, int k, int l, int j) {
  const int sl = k - i - 1, sr_ = j - l - 1;
  if (sl != 0 && sr_ != 0) return z.il(i, k, l, j);
  const int size = sl + sr_, ocb = z.bp(i, j), icb = z.bp(l, k);
  if (size == 0) return z.P->stack[ocb][icb];
  if (size == 1) return z.P->bulge[1] + z.P->stack[ocb][icb];
  return z.P->bulge[size] + z.termau(i, j) + z.termau(l, k);
}

// Tables: record rule 0 = Q, 4 = QM, 5 = Qp (the first rule id of each
// table).  Candidates in the grammar's order (RNAFold.scala:20-64):
//   Q : QM(i,j) [0], Q(i,j-1) [1], Q(i+1,j) [2], Qp(i,j) [3]
//   QM: Q(i,k) + Q(k,j), k = i+3 .. j-3 ascending [4]
//   Qp: hairpin [5], stack [6], interior loops with l outer ascending from
//       max(i+4, j-30) to j-1 and k inner from i+1 to min(i+30, l-3) [7],
//       multiloop QM(i+1, j-1) [8]
template <int NT>
__global__ __launch_bounds__(NT) void rf_finalize(TriArgs a) {
  const TriInst in = a.inst[blockIdx.x];
  const Cells C{a.tab, in.nb, NT_RF, in.tile0};
  const int n = in.n;
  __shared__ int32_t s_head[kRnaHeadWords];
  extern __shared__ uint8_t s_seq[];  // the sequence
  for (int e = threadIdx.x; e < n; e += NT) s_seq[e] = a.codes[in.seq_off + e];
  const ZuF z{s_seq, stage_rna_head(a.rna, s_head), (const DevRna*)a.rna, a.hl_ent, n};
  const int top = C.get(RT_Q, 0, n);
  if (threadIdx.x == 0) a.results[in.out] = top < INV_THR ? top : ZU_EMPTY;
  if (!a.want_trace) return;
  if (top >= INV_THR) {  // no structure: an empty trace
    if (threadIdx.x == 0) a.trace_lens[in.out] = 0;
    return;
  }
  auto rule_of = [&](const TraceRec& cur, auto&& match, int* rule, int* p0, int* p1) -> bool {
    const int i = cur.i, j = cur.j, tabrule = cur.rule, d = j - i;
    const int l0 = fmax_(i + 4, j - 30), nl = fmax_(0, j - 1 - l0 + 1);
    int ncand, target;
    if (tabrule == 0) { target = C.get(RT_Q, i, j); ncand = 4; }
    else if (tabrule == 4) { target = C.get(RT_QM, i, j); ncand = fmax_(0, j - 3 - (i + 3) + 1); }
    else { target = C.get(RT_QP, i, j); ncand = 3 + 30 * nl; }
    auto cand = [&](int c, int* v) -> bool {
      if (tabrule == 0) {
        int x;
        if (c == 0) x = C.get(RT_QM, i, j);
        else if (c == 1) x = C.get(RT_Q, i, j - 1);
        else if (c == 2) x = C.get(RT_Q, i + 1, j);
        else x = C.get(RT_QP, i, j);
        if (x >= INV_THR) return false;
        *v = x; return true;
      }
      if (tabrule == 4) {
        const int k = i + 3 + c;
        const int x = C.get(RT_Q, i, k), y = C.get(RT_Q, k, j);
        if (x >= INV_THR || y >= INV_THR) return false;
        *v = x + y; return true;
      }
      if (c == 0) { if (d < 3) return false; *v = z.hairpin(i, j - 1); return true; }
      if (c == 1) {
        const int x = C.get(RT_QP, i + 1, j - 1);
        if (x >= INV_THR) return false;
        *v = z.sr(i, j - 1) + x; return true;
      }
      if (c < 2 + 30 * nl) {
        const int l = l0 + (c - 2) / 30, k = i + 1 + (c - 2) % 30;
        if (k > l - 3) return false;
        const int x = C.get(RT_QP, k, l);
        if (x >= INV_THR) return false;
        *v = rf_il(z, i, k, l - 1, j - 1) + x; return true;
      }
      const int x = C.get(RT_QM, i + 1, j - 1);
      if (x >= INV_THR) return false;
      *v = x; return true;
    };
    const int c = match(ncand, target, cand);
    if (c < 0) {
      // no candidate: the record is dropped with the whole trace
    } else if (tabrule == 0) {
      *rule = c;
    } else if (tabrule == 4) {
      *rule = 4;
      *p0 = i + 3 + c;
    } else if (c == 0) {
      *rule = 5;
    } else if (c == 1) {
      *rule = 6;
    } else if (c < 2 + 30 * nl) {
      *rule = 7;
      *p0 = i + 1 + (c - 2) % 30;
      *p1 = l0 + (c - 2) / 30;
    } else {
      *rule = 8;
    }
    return c >= 0;
  };
  auto kids = [&](const TraceRec& e, auto&& push) {
    const int i = e.i, j = e.j;
    switch (e.rule) {
      case 0: push(i, j, 4); break;
      case 1: push(i, j - 1, 0); break;
      case 2: push(i + 1, j, 0); break;
      case 3: push(i, j, 5); break;
      case 4: push(i, e.pos[0], 0); push(e.pos[0], j, 0); break;
      case 6: push(i + 1, j - 1, 5); break;
      case 7: push(e.pos[0], e.pos[1], 5); break;
      case 8: push(i + 1, j - 1, 4); break;
      default: break;
    }
  };
  bfs_rounds<NT / 64, (NT >= 1024)>((TraceRec*)a.traces + in.trace_off, in.trace_cap, TraceRec{0, n, 0, {0, 0}}, rule_of, kids,
             a.trace_lens + in.out);
}

// ------------------------------------------------------------------ launch
int launch_finalize(int grammar, const TriArgs& a, int count, int n_max, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (count <= 0) return 0;
  // a few instances: 16 waves each (more backtrace records per round, more
  // st split points in flight); a batch: 4 waves each, several per CU
  const bool big = count <= 32;
  // Zuker: the st column (n + 1 ints) and the sequence; RNAfold: the sequence
  const size_t seq_b = ((size_t)n_max + 4) & ~(size_t)3;
  const size_t lds = grammar == G_ZUKER ? (size_t)(n_max + 1) * sizeof(int) + seq_b : grammar == G_RNAFOLD ? seq_b : 0;
#define DPX_FIN(K)                                                          \
  do {                                                                      \
    if (big) hipLaunchKernelGGL(K<1024>, dim3(count), dim3(1024), lds, st, a); \
    else hipLaunchKernelGGL(K<NTHR>, dim3(count), dim3(NTHR), lds, st, a);     \
  } while (0)
  if (grammar == G_NUSSINOV) DPX_FIN(nu_finalize);
  else if (grammar == G_MATCHAIN) DPX_FIN(mc_finalize);
  else if (grammar == G_RNAFOLD) DPX_FIN(rf_finalize);
  else DPX_FIN(zu_finalize);
#undef DPX_FIN
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace dpx
