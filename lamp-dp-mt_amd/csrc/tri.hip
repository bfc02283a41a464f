// tri.hip -- persistent tiled wavefront kernels for the triangular grammars
// (Nussinov, Zuker MFE, matrix chain) on gfx950.
//
// One workgroup (4 wave64) owns one T x T tile at a time:
//   1. "middle" split points k in blocks strictly between I and J form a
//      tropical (min,+)/(max,+) product of finished tiles; each wave takes
//      every 4th k-block, stages the two operand tiles in LDS and keeps a
//      4x4 register block of cells per lane (2 LDS loads per 16 candidates);
//   2. the rest of each cell (split points inside blocks I and J, and every
//      bounded rule) runs as an anti-diagonal wavefront inside the tile with
//      the dependency frontier (neighbour tiles' halos) staged in LDS;
//   3. the tile is written back and its successors released through
//      device-scope counters; a successor whose two predecessors are done is
//      appended to the shared queue.
// All arithmetic is int32 on the VALU; values only are stored -- the
// backtrace (finalize kernels below) re-derives each winning rule and split
// in the reference's evaluation order, which reproduces the reference's
// back-pointers exactly because its strict comparisons keep the first
// candidate reaching the optimum (CodeGen.scala:122-131).
#include <hip/hip_runtime.h>

#include "engine.h"
#include "rna_model.h"

namespace dpx {

typedef struct { int32_t i, j, rule, pos[2]; } TraceRec;  // == dpx_trace_t

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }

// canonical pair code of bases (x,y) (librna_impl.h:106-114): CG1 GC2 GU3 UG4 AU5 UA6, else 7
constexpr unsigned long long bp_word(int lo) {
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
__device__ __forceinline__ int bpc(int x, int y) {
  const int idx = x * 5 + y;
  const unsigned long long w = idx < 16 ? bp_word(0) : bp_word(16);
  return (int)((w >> (4 * (idx & 15))) & 15);
}

__device__ __forceinline__ int32_t* rec_ptr(const TriArgs& a, const TriInst& in, int NTv, int I, int J) {
  return a.tab + (size_t)(in.tile0 + tile_index(in.nb, I, J)) * (size_t)(NTv * T * T);
}

// LDS ordering inside one wave: all lanes' LDS writes land before later reads.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
}

// ------------------------------------------------------------ middle product
// acc[t][u] (cells (IT+r0+t, JT+c0+u)) = min/max over k in blocks I<K<J of
// L(i,k) (+) R(k,j) (+ extra for the matrix chain).
template <bool MAXP, bool CHAIN>
__device__ void middle_product(const TriArgs& a, const TriInst& in, int NTv, int LT, int RT, int I, int J,
                               int* stage, int acc[4][4]) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r0 = (lane >> 3) * 4, c0 = (lane & 7) * 4;
  int* sL = stage + w * (T * (T + 1) + T * T + T);  // [r][k], row stride T+1
  int* sR = sL + T * (T + 1);                       // [k][c]
  int* sK = sR + T * T;                             // chain: rows[k]
  uint32_t alpha[4][4];
  if (CHAIN) {
    const int32_t* rows = a.dims + in.seq_off;
    const int32_t* cols = rows + in.n;
    for (int t = 0; t < 4; ++t)
      for (int u = 0; u < 4; ++u) {
        int i = I * T + r0 + t, j = J * T + c0 + u;
        alpha[t][u] = (i < in.n && j >= 1 && j <= in.n) ? (uint32_t)rows[i] * (uint32_t)cols[j - 1] : 0u;
      }
  }
  for (int K = I + 1 + w; K < J; K += NWAVE) {
    const int* gL = rec_ptr(a, in, NTv, I, K) + LT * T * T;
    const int* gR = rec_ptr(a, in, NTv, K, J) + RT * T * T;
    for (int q = 0; q < 4; ++q) {
      const int e = (q * 64 + lane) * 4;
      const int4 v = *(const int4*)(gL + e);
      const int r = e / T, k = e % T;
      sL[r * (T + 1) + k + 0] = v.x;
      sL[r * (T + 1) + k + 1] = v.y;
      sL[r * (T + 1) + k + 2] = v.z;
      sL[r * (T + 1) + k + 3] = v.w;
      *(int4*)(sR + e) = *(const int4*)(gR + e);
    }
    if (CHAIN && lane < T) {
      const int kk = K * T + lane;
      sK[lane] = kk < in.n ? a.dims[in.seq_off + kk] : 0;
    }
    wave_lds_sync();
#pragma unroll 2
    for (int k = 0; k < T; ++k) {
      int av[4];
      for (int t = 0; t < 4; ++t) av[t] = sL[(r0 + t) * (T + 1) + k];
      const int4 b = *(const int4*)(sR + k * T + c0);
      const int bv[4] = {b.x, b.y, b.z, b.w};
      uint32_t rk = CHAIN ? (uint32_t)sK[k] : 0u;
      for (int t = 0; t < 4; ++t)
        for (int u = 0; u < 4; ++u) {
          int c;
          if (CHAIN) c = (int)((uint32_t)av[t] + (uint32_t)bv[u] + alpha[t][u] * rk);
          else c = av[t] + bv[u];
          acc[t][u] = MAXP ? imax(acc[t][u], c) : imin(acc[t][u], c);
        }
    }
    wave_lds_sync();
  }
}

// ============================================================ queue driver
// Wave-uniform 64-bit value (lane 0's).
__device__ __forceinline__ unsigned long long uniform64(unsigned long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// Claim the next queue slot and wait for its tile.  Called by all of wave 0
// with uniform control flow (no lane-divergent loop may sit between two
// workgroup barriers: the compiler may then let lanes 1..63 reach the next
// barrier before lane 0's loop has run).
__device__ __forceinline__ unsigned long long claim_item(const TriArgs& a) {
  int slot = 0;
  if (threadIdx.x == 0) slot = __hip_atomic_fetch_add(&a.ctrl[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  slot = __builtin_amdgcn_readfirstlane(slot);
  if (slot >= a.total_tiles) return ~0ull;
  // Poll with relaxed loads (an acquire per poll would invalidate this XCD's
  // L2 every time), then acquire once.  Bounded wait (a few seconds): a
  // scheduling bug ends the kernel with an error flag instead of a hang.
  unsigned long long it = uniform64(__hip_atomic_load(&a.queue[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  for (int spin = 0; it == ~0ull; ++spin) {
    if (spin > (1 << 23)) {
      if (threadIdx.x == 0) __hip_atomic_store(&a.ctrl[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return ~0ull;
    }
    __builtin_amdgcn_s_sleep(4);
    it = uniform64(__hip_atomic_load(&a.queue[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return it;
}

template <class G>
__global__ __launch_bounds__(NTHR) void tri_fill(TriArgs a) {
  __shared__ typename G::Smem sm;
  __shared__ unsigned long long s_item;
  bool live = true;
  while (live) {
    if (threadIdx.x < 64) {
      const unsigned long long it = claim_item(a);
      if (threadIdx.x == 0) s_item = it;
    }
    __syncthreads();
    const unsigned long long it = s_item;
    __syncthreads();
    if (it == ~0ull) {
      live = false;
    } else {
      if (threadIdx.x >= 64) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      const int q = (int)(it >> 32), I = (int)((it >> 16) & 0xffff), J = (int)(it & 0xffff);
      const TriInst in = a.inst[q];
      G::tile(a, in, I, J, sm);
      // every wave's stores have completed before the barrier; thread 0's
      // release RMW below then publishes them at device scope
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
      __syncthreads();
      if (threadIdx.x == 0) {
        // successors (I-1,J) and (I,J+1): queued by whichever predecessor finishes second
        for (int s2 = 0; s2 < 2; ++s2) {
          const int I2 = s2 == 0 ? I - 1 : I, J2 = s2 == 0 ? J : J + 1;
          if (I2 < 0 || J2 >= in.nb) continue;
          const int g = (int)(in.tile0 + tile_index(in.nb, I2, J2));
          const int old = __hip_atomic_fetch_add(&a.dep[g], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
          if (old == 1) {
            const int pos = __hip_atomic_fetch_add(&a.ctrl[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&a.queue[pos], pack_item(q, I2, J2), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
      }
      __syncthreads();
    }
  }
}

// Queue reset: graph-capturable (no memset nodes).
__global__ void tri_reset(TriArgs a, const unsigned long long* seeds, int nseed) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < a.total_tiles) {
    a.queue[t] = t < nseed ? seeds[t] : ~0ull;
    a.dep[t] = 0;
  }
  if (t == 0) { a.ctrl[0] = 0; a.ctrl[1] = nseed; a.ctrl[2] = 0; }
}

// ============================================================== Nussinov
// s(i,j) = max{0 [i==j]; s(i+1,j); s(i,j-1); s(i+1,j-1)+1 [pair(i,j-1), i+2+m<=j];
//             max_{i<k<j} s(i,k)+s(k,j)}      (bench2/nu_kern.h:23-45)
__device__ __forceinline__ bool nu_pair(int x, int y) {
  // au ua gu ug cg gc (Nussinov.scala:50-54)
  const int b = bpc(x, y);
  return b != 7;
}

struct GNu {
  struct Smem {
    union {
      int stage[NWAVE * (T * (T + 1) + T * T + T)];
      struct {
        int S[T + 1][T + 2];  // origin (IT, JT-1)
        int L[T][T + 1];      // tile (I,I) [r][k]
        int R[T][T + 1];      // tile (J,J) [k][c]
      } w;
    } u;
    int X[T][T + 1];
  };

  __device__ static void tile(const TriArgs& a, const TriInst& in, int I, int J, Smem& sm) {
    const int tid = threadIdx.x, IT = I * T, JT = J * T, n = in.n;
    const uint8_t* s = a.codes + in.seq_off;
    int acc[4][4];
    for (int t = 0; t < 4; ++t)
      for (int u = 0; u < 4; ++u) acc[t][u] = NEG;
    for (int e = tid; e < T * (T + 1); e += NTHR) (&sm.X[0][0])[e] = NEG;
    if (J - I >= 2) middle_product<true, false>(a, in, NT_NU, 0, 0, I, J, sm.u.stage, acc);
    __syncthreads();
    {
      const int lane = tid & 63, r0 = (lane >> 3) * 4, c0 = (lane & 7) * 4;
      if (J - I >= 2)
        for (int t = 0; t < 4; ++t)
          for (int u = 0; u < 4; ++u) atomicMax(&sm.X[r0 + t][c0 + u], acc[t][u]);
    }
    // stage the frontier
    for (int e = tid; e < (T + 1) * (T + 1); e += NTHR) {
      const int rr = e / (T + 1), cc = e % (T + 1);
      const int i = IT + rr, j = JT - 1 + cc;
      int v = NEG;
      const bool cur = rr < T && cc >= 1;
      if (!cur && i <= j && j <= n && j >= 0) {
        const int I2 = i / T, J2 = j / T;
        v = rec_ptr(a, in, NT_NU, I2, J2)[(i % T) * T + (j % T)];
      }
      sm.u.w.S[rr][cc] = v;
    }
    if (I < J) {
      const int* gL = rec_ptr(a, in, NT_NU, I, I);
      const int* gR = rec_ptr(a, in, NT_NU, J, J);
      for (int e = tid; e < T * T; e += NTHR) {
        sm.u.w.L[e / T][e % T] = gL[e];
        sm.u.w.R[e / T][e % T] = gR[e];
      }
    }
    __syncthreads();
    const int dlo = I == J ? 0 : -(T - 1);
    for (int ld = dlo; ld <= T - 1; ++ld) {
      const int rmin = ld < 0 ? -ld : 0, ncell = T - (ld < 0 ? -ld : ld);
      // phase A: split points inside blocks I and J, 4 threads per cell
      if (tid < ncell * 4) {
        const int p = tid >> 2, part = tid & 3, r = rmin + p, c = r + ld, i = IT + r, j = JT + c;
        if (i < j && j <= n) {
          int k1lo, k1hi, k2lo, k2hi;
          if (I == J) { k1lo = i + 1; k1hi = j - 1; k2lo = 1; k2hi = 0; }
          else { k1lo = i + 1; k1hi = IT + T - 1; k2lo = JT; k2hi = j - 1; }
          const int n1 = imax(0, k1hi - k1lo + 1), n2 = imax(0, k2hi - k2lo + 1), len = n1 + n2;
          const int b0 = part * len / 4, b1 = (part + 1) * len / 4;
          int best = NEG;
          for (int x = b0; x < b1; ++x) {
            int v;
            if (x < n1) {
              const int k = k1lo + x;
              const int left = I == J ? sm.u.w.S[r][k - JT + 1] : sm.u.w.L[r][k - IT];
              v = left + sm.u.w.S[k - IT][c + 1];
            } else {
              const int k = k2lo + (x - n1);
              v = sm.u.w.S[r][k - JT + 1] + sm.u.w.R[k - JT][c];
            }
            best = imax(best, v);
          }
          if (best > NEG) atomicMax(&sm.X[r][c], best);
        }
      }
      __syncthreads();
      // phase B: one thread per cell
      if (tid < ncell) {
        const int r = rmin + tid, c = r + ld, i = IT + r, j = JT + c;
        int v = NEG;
        if (i <= j && j <= n) {
          v = sm.X[r][c];
          if (i == j) v = imax(v, 0);
          if (i + 1 <= j) { v = imax(v, sm.u.w.S[r + 1][c + 1]); v = imax(v, sm.u.w.S[r][c]); }
          if (i + 2 + a.min_loop <= j && nu_pair(s[i], s[j - 1])) v = imax(v, sm.u.w.S[r + 1][c] + 1);
        }
        sm.u.w.S[r][c + 1] = v;
      }
      __syncthreads();
    }
    int* g = rec_ptr(a, in, NT_NU, I, J);
    for (int e = tid; e < T * T; e += NTHR) g[e] = sm.u.w.S[e / T][e % T + 1];
  }
};

// ============================================================ matrix chain
// M(i,j) = min_{i<k<j} M(i,k)+M(k,j)+rows[i]*rows[k]*cols[j-1]  (int32 wrap),
// M(i,i+1) = 0   (Benchmarks.scala:18-27)
struct GMc {
  struct Smem {
    union {
      int stage[NWAVE * (T * (T + 1) + T * T + T)];
      struct {
        int S[T][T + 1];      // current tile
        int L[T][T + 1];      // tile (I,I)
        int R[T][T + 1];      // tile (J,J)
        int rows[2 * T];      // rows[IT..IT+T) , rows[JT..JT+T)
      } w;
    } u;
    int X[T][T + 1];
  };

  __device__ static void tile(const TriArgs& a, const TriInst& in, int I, int J, Smem& sm) {
    const int tid = threadIdx.x, IT = I * T, JT = J * T, n = in.n;
    const int32_t* rows = a.dims + in.seq_off;
    const int32_t* cols = rows + n;
    int acc[4][4];
    for (int t = 0; t < 4; ++t)
      for (int u = 0; u < 4; ++u) acc[t][u] = 0x7fffffff;
    for (int e = tid; e < T * (T + 1); e += NTHR) (&sm.X[0][0])[e] = 0x7fffffff;
    if (J - I >= 2) middle_product<false, true>(a, in, NT_MC, 0, 0, I, J, sm.u.stage, acc);
    __syncthreads();
    {
      const int lane = tid & 63, r0 = (lane >> 3) * 4, c0 = (lane & 7) * 4;
      if (J - I >= 2)
        for (int t = 0; t < 4; ++t)
          for (int u = 0; u < 4; ++u) atomicMin(&sm.X[r0 + t][c0 + u], acc[t][u]);
    }
    for (int e = tid; e < T * (T + 1); e += NTHR) (&sm.u.w.S[0][0])[e] = 0;
    if (I < J) {
      const int* gL = rec_ptr(a, in, NT_MC, I, I);
      const int* gR = rec_ptr(a, in, NT_MC, J, J);
      for (int e = tid; e < T * T; e += NTHR) {
        sm.u.w.L[e / T][e % T] = gL[e];
        sm.u.w.R[e / T][e % T] = gR[e];
      }
    }
    if (tid < 2 * T) {
      const int k = tid < T ? IT + tid : JT + tid - T;
      sm.u.w.rows[tid] = k < n ? rows[k] : 0;
    }
    __syncthreads();
    const int dlo = I == J ? 0 : -(T - 1);
    for (int ld = dlo; ld <= T - 1; ++ld) {
      const int rmin = ld < 0 ? -ld : 0, ncell = T - (ld < 0 ? -ld : ld);
      if (tid < ncell * 4) {
        const int p = tid >> 2, part = tid & 3, r = rmin + p, c = r + ld, i = IT + r, j = JT + c;
        if (i + 1 < j && j <= n) {
          int k1lo, k1hi, k2lo, k2hi;
          if (I == J) { k1lo = i + 1; k1hi = j - 1; k2lo = 1; k2hi = 0; }
          else { k1lo = i + 1; k1hi = IT + T - 1; k2lo = JT; k2hi = j - 1; }
          const int n1 = imax(0, k1hi - k1lo + 1), n2 = imax(0, k2hi - k2lo + 1), len = n1 + n2;
          const int b0 = part * len / 4, b1 = (part + 1) * len / 4;
          const uint32_t alpha = (uint32_t)rows[i] * (uint32_t)cols[j - 1];
          int best = 0x7fffffff;
          for (int x = b0; x < b1; ++x) {
            uint32_t v;
            if (x < n1) {
              const int k = k1lo + x;
              const int left = I == J ? sm.u.w.S[r][k - JT] : sm.u.w.L[r][k - IT];
              v = (uint32_t)left + (uint32_t)sm.u.w.S[k - IT][c] + alpha * (uint32_t)sm.u.w.rows[k - IT];
            } else {
              const int k = k2lo + (x - n1);
              v = (uint32_t)sm.u.w.S[r][k - JT] + (uint32_t)sm.u.w.R[k - JT][c] + alpha * (uint32_t)sm.u.w.rows[T + k - JT];
            }
            best = imin(best, (int)v);
          }
          if (b1 > b0) atomicMin(&sm.X[r][c], best);
        }
      }
      __syncthreads();
      if (tid < ncell) {
        const int r = rmin + tid, c = r + ld, i = IT + r, j = JT + c;
        int v = 0;
        if (i + 1 == j && j <= n) v = 0;
        else if (i + 1 < j && j <= n) v = sm.X[r][c];
        sm.u.w.S[r][c] = v;
      }
      __syncthreads();
    }
    int* g = rec_ptr(a, in, NT_MC, I, J);
    for (int e = tid; e < T * T; e += NTHR) g[e] = sm.u.w.S[e / T][e % T];
  }
};

// ================================================================ Zuker
// Tables per tile: cl, U, ml, ml1, R11 (zu_kern.h:92-184, values only):
//   U(i,k)   = cl(i,k) + ext_mismatch(i,k-1) + termau(i,k-1) + ML_intern  (fun8+fun9)
//   X(i,j)   = min_{k} U(i,k) + ml1(k,j)              (rules 7 and 9 share it)
//   ml(i,j)  = min{ ml(i+1,j) [j-i>=15]; X }
//   ml1(i,j) = min{ ml1(i+1,j) [j-i>=8]; X; U(i,j); R11(i,j) }
//   R11(i,j) = min_{i+7<=k<j} U(i,k) = min{R11(i,j-1), U(i,j-1)}     (rule 11)
// INV marks "no valid candidate", i.e. the reference's 999999.
struct ZuCtx {
  const uint8_t* s;
  const DevRna* P;
  const int32_t* hl;
  int n;
  __device__ int S(int k) const { return s[k]; }
  __device__ int bp(int x, int y) const { return bpc(S(x), S(y)); }
  __device__ int termau(int x, int y) const {
    const int a = S(x), b = S(y);
    return ((a == 3 && b == 2) || (a == 2 && b == 3)) ? 0 : P->term_au;
  }
  __device__ int sr(int i, int j) const { return P->stack[bp(i, j)][bp(j - 1, i + 1)]; }
  // inner mismatch of cell (l,k) (pair l..k-1) as an interior-loop closing pair
  __device__ int inner_I(int l, int k) const { return P->mmI[bpc(S(k - 1), S(l))][S(k)][S(l - 1)]; }
  __device__ int ext_mm(int i, int j) const {
    const int b = bp(i, j);
    if (i > 0 && j + 1 < n) return P->mmExt[b][S(i - 1)][S(j + 1)];
    if (i > 0) return P->d5[b][S(i - 1)];
    if (j + 1 < n) return P->d3[b][S(j + 1)];
    return 0;
  }
  // hairpin closed by (i,j) (librna_impl.h:181-193)
  __device__ int hairpin(int i, int j) const {
    const int size = j - i - 1;
    if (size < 3) return 600;
    int cls = size == 3 ? 0 : size == 4 ? 1 : size == 6 ? 2 : -1;
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
  // interior loop (i,j) closing (k,l), (librna_impl.h:218-265) for sl,sr >= 1
  __device__ int il(int i, int k, int l, int j) const {
    const int sl = k - i - 1, sr_ = j - l - 1;
    const int ocb = bp(i, j), icb = bp(l, k);
    if (sl == 1) {
      if (sr_ == 1) return P->int11[ocb][bp(j - 2, i + 2)][S(i + 1)][S(j - 1)];
      if (sr_ == 2) return P->int21[ocb][bp(j - 3, i + 2)][S(i + 1)][S(j - 2)][S(j - 1)];
      return P->F[sl][sr_] + P->mm1nI[ocb][S(i + 1)][S(j - 1)] + P->mm1nI[icb][S(l + 1)][S(k - 1)];
    }
    if (sl == 2) {
      if (sr_ == 1) return P->int21[bp(j - 2, i + 3)][ocb][S(j - 1)][S(i + 1)][S(i + 2)];
      if (sr_ == 2) return P->int22[ocb][bp(j - 3, i + 3)][S(i + 1)][S(i + 2)][S(j - 2)][S(j - 1)];
      if (sr_ == 3) return P->il5 + P->ninio0 + P->mm23I[ocb][S(i + 1)][S(j - 1)] + P->mm23I[icb][S(l + 1)][S(k - 1)];
    } else if (sl == 3 && sr_ == 2) {
      return P->il5 + P->ninio0 + P->mm23I[ocb][S(i + 1)][S(j - 1)] + P->mm23I[icb][S(l + 1)][S(k - 1)];
    } else if (sr_ == 1) {
      return P->F[sl][sr_] + P->mm1nI[ocb][S(i + 1)][S(j - 1)] + P->mm1nI[icb][S(l + 1)][S(k - 1)];
    }
    return P->F[sl][sr_] + P->mmI[ocb][S(i + 1)][S(j - 1)] + P->mmI[icb][S(l + 1)][S(k - 1)];
  }
  // bulge terms of rules 2 and 3 (bl_energy/br_energy, librna_impl.h:199-215)
  __device__ int bulge_l(int i, int j, int k) const {  // cl(k,j-2) inside, size k-i-2
    const int size = k - i - 2;
    if (size == 1) return P->bulge[1] + P->stack[bp(i + 1, j - 2)][bp(j - 3, k)];
    return P->bulge[size] + termau(i + 1, j - 2) + termau(j - 3, k);
  }
  __device__ int bulge_r(int i, int j, int k) const {  // cl(i+2,k) inside, size j-k-2
    const int size = j - k - 2;
    if (size == 1) return P->bulge[1] + P->stack[bp(i + 1, j - 2)][bp(k - 1, i + 2)];
    return P->bulge[size] + termau(i + 1, j - 2) + termau(k - 1, i + 2);
  }
  __device__ bool stackpair(int i, int j) const {  // fun6 (zu.c:35)
    return i + 4 <= j && bp(i, j - 1) != 7 && bp(i + 1, j - 2) != 7;
  }
  __device__ int rule5(int i, int j, int mlv) const {
    return P->ml_closing + P->ml_intern + mlv + termau(i + 1, j - 2) + P->mmM[bp(j - 2, i + 1)][S(j - 3)][S(i + 2)];
  }
};

struct GZu {
  static constexpr int BR = 64, BC = 64;   // B region: rows IT..IT+63, cols JT-32..JT+31
  struct Smem {
    union {
      int stage[NWAVE * (T * (T + 1) + T * T + T)];
      struct {
        int B[BR][BC + 1];     // cl + inner_I, origin (IT, JT-32)
        int U[T][T + 2];       // origin (IT, JT-1)
        int ML[T + 2][T + 3];  // origin (IT, JT-2)
        int ML1[T + 1][T + 1]; // origin (IT, JT)
        int R11[T][T + 2];     // origin (IT, JT-1)
        int LU[T][T + 1];      // U tile (I,I)
        int RM1[T][T + 1];     // ml1 tile (J,J)
      } w;
    } u;
    int X[T][T + 1];
    int16_t F[31][32];
    int acc[T];
    int list[2][T];
    int nlist[2];
  };

  __device__ static int cell_valid_ml(int d) { return d >= 12; }

  __device__ static void build_list(const ZuCtx& z, int IT, int JT, int ld, int* list, int* nlist, int lane) {
    // cells of local diagonal ld whose cl can use bulges / interior loops
    const int rmin = ld < 0 ? -ld : 0, ncell = T - (ld < 0 ? -ld : ld);
    bool ok = false;
    if (lane < ncell) {
      const int r = rmin + lane, i = IT + r, j = JT + r + ld;
      ok = j <= z.n && j - i >= 12 && z.stackpair(i, j);
    }
    const unsigned long long m = __ballot(ok);
    if (ok) list[__popcll(m & ((1ull << lane) - 1))] = lane;
    if (lane == 0) *nlist = __popcll(m);
  }

  __device__ static void tile(const TriArgs& a, const TriInst& in, int I, int J, Smem& sm) {
    const int tid = threadIdx.x, IT = I * T, JT = J * T, n = in.n, nb = in.nb;
    ZuCtx z{a.codes + in.seq_off, (const DevRna*)a.rna, a.hl_ent, n};
    const DevRna* P = z.P;
    int acc[4][4];
    for (int t = 0; t < 4; ++t)
      for (int u = 0; u < 4; ++u) acc[t][u] = INV;
    for (int e = tid; e < T * (T + 1); e += NTHR) (&sm.X[0][0])[e] = INV;
    for (int e = tid; e < 31 * 31; e += NTHR) sm.F[e / 31][e % 31] = P->F[e / 31][e % 31];
    if (tid < T) sm.acc[tid] = INV;
    if (J - I >= 2) middle_product<false, false>(a, in, NT_ZU, ZT_U, ZT_ML1, I, J, sm.u.stage, acc);
    __syncthreads();
    {
      const int lane = tid & 63, r0 = (lane >> 3) * 4, c0 = (lane & 7) * 4;
      if (J - I >= 2)
        for (int t = 0; t < 4; ++t)
          for (int u = 0; u < 4; ++u)
            if (acc[t][u] < INV_THR) atomicMin(&sm.X[r0 + t][c0 + u], acc[t][u]);
    }
    // ---- stage the frontier (values of finished neighbour tiles) ----
    auto fetch = [&](int tab, int i, int j) -> int {
      if (i < 0 || j < 0 || i > j || j > n) return INV;
      const int I2 = i / T, J2 = j / T;
      if (I2 >= nb || J2 >= nb) return INV;
      return rec_ptr(a, in, NT_ZU, I2, J2)[tab * T * T + (i % T) * T + (j % T)];
    };
    for (int e = tid; e < BR * BC; e += NTHR) {
      const int rr = e / BC, cc = e % BC, l = IT + rr, k = JT - 32 + cc;
      int v = INV;
      if (!(rr < T && cc >= 32)) {
        const int cl = fetch(ZT_CL, l, k);
        if (cl < INV_THR && l >= 1 && k <= n - 1) v = cl + z.inner_I(l, k);
      }
      sm.u.w.B[rr][cc] = v;
    }
    for (int e = tid; e < T * (T + 1); e += NTHR) {
      const int rr = e / (T + 1), cc = e % (T + 1);
      sm.u.w.U[rr][cc] = cc == 0 ? fetch(ZT_U, IT + rr, JT - 1) : INV;
      sm.u.w.R11[rr][cc] = cc == 0 ? fetch(ZT_R11, IT + rr, JT - 1) : INV;
    }
    for (int e = tid; e < (T + 2) * (T + 2); e += NTHR) {
      const int rr = e / (T + 2), cc = e % (T + 2);
      sm.u.w.ML[rr][cc] = (rr >= T || cc < 2) ? fetch(ZT_ML, IT + rr, JT - 2 + cc) : INV;
    }
    for (int e = tid; e < (T + 1) * T; e += NTHR) {
      const int rr = e / T, cc = e % T;
      sm.u.w.ML1[rr][cc] = rr == T ? fetch(ZT_ML1, IT + T, JT + cc) : INV;
    }
    if (I < J) {
      const int* gL = rec_ptr(a, in, NT_ZU, I, I) + ZT_U * T * T;
      const int* gR = rec_ptr(a, in, NT_ZU, J, J) + ZT_ML1 * T * T;
      for (int e = tid; e < T * T; e += NTHR) {
        sm.u.w.LU[e / T][e % T] = gL[e];
        sm.u.w.RM1[e / T][e % T] = gR[e];
      }
    }
    const int dlo = I == J ? 0 : -(T - 1);
    if (tid >= 64 && tid < 128) build_list(z, IT, JT, dlo, sm.list[0], &sm.nlist[0], tid - 64);
    __syncthreads();
    int32_t* gcl = rec_ptr(a, in, NT_ZU, I, J) + ZT_CL * T * T;
    int buf = 0;
    for (int ld = dlo; ld <= T - 1; ++ld, buf ^= 1) {
      const int rmin = ld < 0 ? -ld : 0, ncell = T - (ld < 0 ? -ld : ld);
      const int nvalid = sm.nlist[buf];
      const int total = ncell * 4 + nvalid * 34;
      // ---------------- phase A ----------------
      for (int w = tid; w < total; w += NTHR) {
        if (w < ncell * 4) {
          // split points of X inside blocks I and J
          const int p = w >> 2, part = w & 3, r = rmin + p, c = r + ld, i = IT + r, j = JT + c;
          if (j > n || j - i < 14) continue;
          int k1lo, k1hi, k2lo, k2hi;
          if (I == J) { k1lo = i + 7; k1hi = j - 7; k2lo = 1; k2hi = 0; }
          else { k1lo = i + 7; k1hi = imin(IT + T - 1, j - 7); k2lo = imax(JT, i + 7); k2hi = j - 7; }
          const int n1 = imax(0, k1hi - k1lo + 1), n2 = imax(0, k2hi - k2lo + 1), len = n1 + n2;
          const int b0 = part * len / 4, b1 = (part + 1) * len / 4;
          int best = INV;
          for (int x = b0; x < b1; ++x) {
            int v;
            if (x < n1) {
              const int k = k1lo + x;
              const int left = I == J ? sm.u.w.U[r][k - JT + 1] : sm.u.w.LU[r][k - IT];
              v = left + sm.u.w.ML1[k - IT][c];
            } else {
              const int k = k2lo + (x - n1);
              v = sm.u.w.U[r][k - JT + 1] + sm.u.w.RM1[k - JT][c];
            }
            best = imin(best, v);
          }
          if (best < INV_THR) atomicMin(&sm.X[r][c], best);
        } else {
          const int w2 = w - ncell * 4, v = w2 / 34, slot = w2 % 34;
          const int p = sm.list[buf][v], r = rmin + p, c = r + ld, i = IT + r, j = JT + c, d = j - i;
          const int srv = z.sr(i, j - 1);
          int best = INV;
          if (slot < 30) {
            // interior loops with left size a = slot+1 (rule 4)
            const int aa = slot + 1, bmax = imin(30, d - 11 - aa);
            if (bmax >= 1) {
              const int lrow = r + 2 + aa, l = i + 2 + aa;
              int gen = INV;
              if (aa >= 2) {
                const int bst = aa == 2 ? 4 : aa == 3 ? 3 : 2;
                for (int b = bst; b <= bmax; ++b) gen = imin(gen, sm.u.w.B[lrow][c + 30 - b] + sm.F[aa][b]);
                if (gen < INV_THR) gen += P->mmI[z.bp(i + 1, j - 2)][z.S(i + 2)][z.S(j - 3)];
              }
              best = gen;
              // the special sizes: a==1 row, b==1 column, (2,2) (2,3) (3,2)
              const int bsp = aa == 1 ? bmax : aa == 2 ? imin(3, bmax) : aa == 3 ? imin(2, bmax) : 1;
              for (int b = 1; b <= bsp; ++b) {
                const int k = j - 2 - b;
                const int Bv = sm.u.w.B[lrow][c + 30 - b];
                if (Bv >= INV_THR) continue;
                const int cl = Bv - z.inner_I(l, k);
                best = imin(best, cl + z.il(i + 1, l, k - 1, j - 2));
              }
            }
          } else if (slot < 32) {
            // left bulges (rule 2), half of the k range each
            if (d >= 12) {
              const int klo = i + 3, khi = imin(j - 9, i + 32), len = khi - klo + 1;
              const int h = slot - 30, x0 = klo + h * len / 2, x1 = klo + (h + 1) * len / 2;
              for (int k = x0; k < x1; ++k) {
                const int Bv = sm.u.w.B[k - IT][c + 30];
                if (Bv >= INV_THR) continue;
                best = imin(best, Bv - z.inner_I(k, j - 2) + z.bulge_l(i, j, k));
              }
            }
          } else {
            // right bulges (rule 3)
            if (d >= 12) {
              const int klo = imax(i + 9, j - 32), khi = j - 3, len = khi - klo + 1;
              const int h = slot - 32, x0 = klo + h * len / 2, x1 = klo + (h + 1) * len / 2;
              for (int k = x0; k < x1; ++k) {
                const int Bv = sm.u.w.B[r + 2][k - JT + 32];
                if (Bv >= INV_THR) continue;
                best = imin(best, Bv - z.inner_I(i + 2, k) + z.bulge_r(i, j, k));
              }
            }
          }
          if (best < INV_THR) atomicMin(&sm.acc[p], best + srv);
        }
      }
      __syncthreads();
      // ---------------- phase B: one thread per cell ----------------
      if (tid < T) {
        const int p = tid;
        if (p < ncell) {
          const int r = rmin + p, c = r + ld, i = IT + r, j = JT + c, d = j - i;
          int cl = INV, U = INV, ml = INV, ml1 = INV, R11 = INV, Bv = INV;
          if (i <= j && j <= n) {
            if (d >= 7 && z.stackpair(i, j)) {
              const int srv = z.sr(i, j - 1);
              cl = sm.acc[p];
              if (d >= 9) {
                const int Bi = sm.u.w.B[r + 1][c + 31];
                if (Bi < INV_THR) cl = imin(cl, Bi - z.inner_I(i + 1, j - 1) + srv);
              }
              cl = imin(cl, z.hairpin(i + 1, j - 2) + srv);
              if (d >= 18) {
                const int mlv = sm.u.w.ML[r + 2][c];
                if (mlv < INV_THR) cl = imin(cl, z.rule5(i, j, mlv) + srv);
              }
              U = cl + z.ext_mm(i, j - 1) + z.termau(i, j - 1) + P->ml_intern;
              if (i >= 1 && j <= n - 1) Bv = cl + z.inner_I(i, j);
            }
            int X = sm.X[r][c];
            if (X >= INV_THR) X = INV;
            ml = X;
            if (d >= 15) ml = imin(ml, sm.u.w.ML[r + 1][c + 2]);
            if (d >= 8) {
              const int Rp = sm.u.w.R11[r][c], Up = sm.u.w.U[r][c];
              R11 = (d >= 9) ? imin(Rp, Up) : Up;
            }
            ml1 = imin(imin(X, U), R11);
            if (d >= 8) ml1 = imin(ml1, sm.u.w.ML1[r + 1][c]);
            if (ml >= INV_THR) ml = INV;
            if (ml1 >= INV_THR) ml1 = INV;
            if (R11 >= INV_THR) R11 = INV;
          }
          sm.u.w.B[r][c + 32] = Bv;
          sm.u.w.U[r][c + 1] = U;
          sm.u.w.ML[r][c + 2] = ml;
          sm.u.w.ML1[r][c] = ml1;
          sm.u.w.R11[r][c + 1] = R11;
          gcl[r * T + c] = cl;
        }
        sm.acc[p] = INV;
      } else if (tid >= 64 && tid < 128 && ld + 1 <= T - 1) {
        build_list(z, IT, JT, ld + 1, sm.list[buf ^ 1], &sm.nlist[buf ^ 1], tid - 64);
      }
      __syncthreads();
    }
    int32_t* g = rec_ptr(a, in, NT_ZU, I, J);
    for (int e = tid; e < T * T; e += NTHR) {
      const int r = e / T, c = e % T;
      g[ZT_U * T * T + e] = sm.u.w.U[r][c + 1];
      g[ZT_ML * T * T + e] = sm.u.w.ML[r][c + 2];
      g[ZT_ML1 * T * T + e] = sm.u.w.ML1[r][c];
      g[ZT_R11 * T * T + e] = sm.u.w.R11[r][c + 1];
    }
  }
};

// ------------------------------------------------------------------ launch
int fill_occupancy(int grammar) {
  int blocks = 0;
  hipError_t e;
  if (grammar == G_NUSSINOV) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, tri_fill<GNu>, NTHR, 0);
  else if (grammar == G_MATCHAIN) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, tri_fill<GMc>, NTHR, 0);
  else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, tri_fill<GZu>, NTHR, 0);
  return e == hipSuccess ? blocks : 0;
}

int launch_fill(int grammar, const TriArgs& a, const unsigned long long* seeds, int nseed, int grid, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(tri_reset, dim3((a.total_tiles + 255) / 256), dim3(256), 0, st, a, seeds, nseed);
  if (grammar == G_NUSSINOV) hipLaunchKernelGGL(tri_fill<GNu>, dim3(grid), dim3(NTHR), 0, st, a);
  else if (grammar == G_MATCHAIN) hipLaunchKernelGGL(tri_fill<GMc>, dim3(grid), dim3(NTHR), 0, st, a);
  else hipLaunchKernelGGL(tri_fill<GZu>, dim3(grid), dim3(NTHR), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace dpx
