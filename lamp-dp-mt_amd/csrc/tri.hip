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

// min over x in [0, T) of a[x] + b[x * bstride], both in LDS.  The loads go
// out 8 pairs at a time ahead of their adds (left alone, the scheduler
// reuses two registers and waits for each load in turn).
__device__ __forceinline__ int lds_minplus8(const int* a, const int* b, int bstride) {
  int best = INV;
#pragma unroll
  for (int x0 = 0; x0 < T; x0 += 8) {
    int av[8], bv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      av[u] = a[x0 + u];
      bv[u] = b[(x0 + u) * bstride];
    }
    __builtin_amdgcn_sched_group_barrier(0x100, 12, 0);  // LDS reads
    __builtin_amdgcn_sched_group_barrier(0x002, 16, 0);  // VALU
#pragma unroll
    for (int u = 0; u < 8; ++u) best = imin(best, av[u] + bv[u]);
  }
  return best;
}
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
// L(i,k) (+) R(k,j) (+ extra for the matrix chain).  Per wave, LDS holds the
// left operand k-major (row r of column k at [k][r ^ mp_swz(k)]: one 128-bit
// read gives a lane its 4 rows at one k, and the XOR keeps the transposing
// writes at 2 lanes per bank), the right operand row-major ([k][c]) and, for
// the chain, rows[k].
constexpr int MP_STAGE = T * T + T * T + T;
__device__ __forceinline__ int mp_swz(int k) { return ((k >> 2) & 7) << 2; }
template <bool MAXP, bool CHAIN>
__device__ void middle_product(const TriArgs& a, const TriInst& in, int NTv, int LT, int RT, int I, int J,
                               int* stage, int acc[4][4], int part = 0, int P = 1) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r0 = (lane >> 3) * 4, c0 = (lane & 7) * 4;
  int* sL = stage + w * MP_STAGE;  // [k][r ^ mp_swz(k)]
  int* sR = sL + T * T;            // [k][c]
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
  // operand tiles of block K travel in registers: those of the next block
  // are loaded while this block's products are computed.  Named registers,
  // not arrays: an array indexed in a loop may be left in scratch memory.
  int4 vl0, vl1, vl2, vl3, vr0, vr1, vr2, vr3;
  auto load = [&](int K) {
    const int4* gL = (const int4*)(rec_ptr(a, in, NTv, I, K) + LT * T * T) + lane;
    const int4* gR = (const int4*)(rec_ptr(a, in, NTv, K, J) + RT * T * T) + lane;
    vl0 = gL[0];
    vr0 = gR[0];
    vl1 = gL[64];
    vr1 = gR[64];
    vl2 = gL[128];
    vr2 = gR[128];
    vl3 = gL[192];
    vr3 = gR[192];
  };
  // element e = (q * 64 + lane) * 4 of both operands into LDS
  auto stage_q = [&](int q, int4 l, int4 rv) {
    const int e = (q * 64 + lane) * 4;
    const int r = e / T, k = e % T, rs = r ^ mp_swz(k);  // k % 4 == 0: one swizzle for k..k+3
    sL[(k + 0) * T + rs] = l.x;
    sL[(k + 1) * T + rs] = l.y;
    sL[(k + 2) * T + rs] = l.z;
    sL[(k + 3) * T + rs] = l.w;
    *(int4*)(sR + e) = rv;
  };
  // blocks K = I + 1 + part + P m; wave w takes every NWAVE-th of them
  const int K0 = I + 1 + part + P * w, KS = P * NWAVE;
  if (K0 < J) load(K0);
  for (int K = K0; K < J; K += KS) {
    stage_q(0, vl0, vr0);
    stage_q(1, vl1, vr1);
    stage_q(2, vl2, vr2);
    stage_q(3, vl3, vr3);
    if (K + KS < J) load(K + KS);
    if (CHAIN && lane < T) {
      const int kk = K * T + lane;
      sK[lane] = kk < in.n ? a.dims[in.seq_off + kk] : 0;
    }
    wave_lds_sync();
#pragma unroll 2
    for (int k = 0; k < T; ++k) {
      const int4 a4 = *(const int4*)(sL + k * T + (r0 ^ mp_swz(k)));  // rows r0..r0+3 in order
      const int av[4] = {a4.x, a4.y, a4.z, a4.w};
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

// The middle product of tile (I,J) reduced into X, which the caller has set
// to MP::NEUTRAL; ends with a workgroup barrier.  PART: the blocks
// K = I + 1 + part + P m only (part task `part` of P).  Otherwise P == 1:
// every split block here; P > 1: the P part tasks' partial results, from
// a.mp.  min/max is associative and the candidates the threshold drops
// (Zuker-like MP::THR) are dropped in either order, so every way gives the
// same X.
template <class MP, bool PART>
__device__ void middle_X(const TriArgs& a, const TriInst& in, int I, int J, int* stage, int (*X)[T + 1], int part,
                         int P) {
  const int tid = threadIdx.x;
  if (J - I < 2) {
    __syncthreads();
    return;
  }
  if (!PART && P > 1) {
    __syncthreads();  // X's initial values are in place
    const int g = (int)(in.tile0 + tile_index(in.nb, I, J));
    const int32_t* src = a.mp + mp_part_offset(a.total_tiles) + (int64_t)a.mp[a.total_tiles + g] * (T * T);
    for (int e = tid; e < T * T; e += NTHR) {
      int v = X[e / T][e % T];
      for (int p = 0; p < P; ++p) {
        const int c = src[(size_t)p * (T * T) + e];
        v = MP::MAXP ? imax(v, c) : imin(v, c);
      }
      X[e / T][e % T] = v;
    }
    __syncthreads();
    return;
  }
  int acc[4][4];
  for (int t = 0; t < 4; ++t)
    for (int u = 0; u < 4; ++u) acc[t][u] = MP::NEUTRAL;
  if (PART) middle_product<MP::MAXP, MP::CHAIN>(a, in, MP::NTv, MP::LT, MP::RT, I, J, stage, acc, part, P);
  else middle_product<MP::MAXP, MP::CHAIN>(a, in, MP::NTv, MP::LT, MP::RT, I, J, stage, acc);
  __syncthreads();
  const int lane = tid & 63, r0 = (lane >> 3) * 4, c0 = (lane & 7) * 4;
  for (int t = 0; t < 4; ++t)
    for (int u = 0; u < 4; ++u) {
      if (MP::THR && acc[t][u] >= INV_THR) continue;
      if (MP::MAXP) atomicMax(&X[r0 + t][c0 + u], acc[t][u]);
      else atomicMin(&X[r0 + t][c0 + u], acc[t][u]);
    }
  __syncthreads();
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
  if (slot >= a.total_items) return ~0ull;
  // Poll with relaxed loads.  Bounded wait (a few seconds): a scheduling bug
  // ends the kernel with an error flag instead of a hang.
  unsigned long long it = uniform64(__hip_atomic_load(&a.queue[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  for (int spin = 0; it == ~0ull; ++spin) {
    if (spin > (1 << 23)) {
      if (threadIdx.x == 0) __hip_atomic_store(&a.ctrl[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return ~0ull;
    }
    __builtin_amdgcn_s_sleep(4);
    it = uniform64(__hip_atomic_load(&a.queue[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
  // Acquire at device scope (invalidates this CU's L1 and this XCD's L2
  // copies of other XCDs' data): the records this tile reads were written
  // back by their producers before the item was published, and no copy of
  // them cached here, from this launch or an earlier one using the same
  // addresses, may be read in their place.
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return it;
}

// A part task of a split middle product: its blocks into X, X to its partial slot.
template <class G>
__device__ void mp_part_task(const TriArgs& a, const TriInst& in, int I, int J, int part, int P, typename G::Smem& sm) {
  for (int e = threadIdx.x; e < T * (T + 1); e += NTHR) (&sm.X[0][0])[e] = G::MP::NEUTRAL;
  middle_X<typename G::MP, true>(a, in, I, J, sm.u.stage, sm.X, part, P);
  const int g = (int)(in.tile0 + tile_index(in.nb, I, J));
  int32_t* dst = a.mp + mp_part_offset(a.total_tiles) + (int64_t)(a.mp[a.total_tiles + g] + part) * (T * T);
  for (int e = threadIdx.x; e < T * T; e += NTHR) dst[e] = sm.X[e / T][e % T];
}

template <class G>
__global__ __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(G::WAVES))) void tri_fill(TriArgs a) {
  __shared__ typename G::Smem sm;
  __shared__ unsigned long long s_item;
  __shared__ int s_fin;
  G::init(a, sm);
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
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      const int q = (int)(it >> 32), I = (int)((it >> 16) & 0xfff), part = (int)((it >> 28) & 0xf),
                J = (int)(it & 0xffff);
      const TriInst in = a.inst[q];
      const int P = a.mp_split ? mp_parts(J - I) : 1;
      if (P > 1) {
        // a part task; the one that completes the count goes on with the tile
        mp_part_task<G>(a, in, I, J, part, P, sm);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the partial is written
        __syncthreads();
        if (threadIdx.x == 0) {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
          const int g = (int)(in.tile0 + tile_index(in.nb, I, J));
          const int old = __hip_atomic_fetch_add(&a.mp[g], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          s_fin = old == P - 1;
          // the other parts' partials, written back before their counts
          if (old == P - 1) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        if (!s_fin) continue;
      }
      G::tile(a, in, I, J, sm, 0, P);
      // every wave's stores have completed before the barrier; thread 0's
      // release RMW below then publishes them at device scope
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
      __syncthreads();
      if (threadIdx.x == 0) {
        // one device-scope release (L2 writeback) covers the whole record
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        // successors (I-1,J) and (I,J+1): queued by whichever predecessor finishes second
        for (int s2 = 0; s2 < 2; ++s2) {
          const int I2 = s2 == 0 ? I - 1 : I, J2 = s2 == 0 ? J : J + 1;
          if (I2 < 0 || J2 >= in.nb) continue;
          const int g = (int)(in.tile0 + tile_index(in.nb, I2, J2));
          const int old = __hip_atomic_fetch_add(&a.dep[g], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (old == 1) {
            const int P2 = a.mp_split ? mp_parts(J2 - I2) : 1;
            const int pos = __hip_atomic_fetch_add(&a.ctrl[1], P2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (int p2 = 0; p2 < P2; ++p2)
              __hip_atomic_store(&a.queue[pos + p2], pack_item(q, I2, J2, p2), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
          }
        }
      }
      __syncthreads();
    }
  }
}

// Special hairpins (hl_energy's strMatch over Triloops / Tetraloops /
// Hexaloops, librna_impl.h:154-193): spec[cls][p] = energy of the special
// loop of size 3, 4 or 6 (cls 0, 1, 2) whose closing pair starts at code
// position p, i.e. the size+2 bases codes[p..], or SPEC_NONE.  One thread
// per position; instances are separated by a 0 code, which no entry holds.
__global__ void tri_spec(TriArgs a) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n_codes) return;
  const DevRna* P = (const DevRna*)a.rna;
  for (int cls = 0; cls < 3; ++cls) {
    const int len = (cls == 0 ? 3 : cls == 1 ? 4 : 6) + 2;
    int v = SPEC_NONE;
    if (p + len <= a.n_codes) {
      int key = 0;
      for (int m = 0; m < len; ++m) key = key * 8 + a.codes[p + m];
      for (int e = 0; e < P->n_loop[cls]; ++e)
        if (P->loop_key[cls][e] == key) {
          v = P->loop_e[cls][e];
          break;
        }
    }
    a.spec[cls * a.n_codes + p] = (int16_t)v;
  }
}

// Queue reset: graph-capturable (no memset nodes).
__global__ void tri_reset(TriArgs a, const unsigned long long* seeds, int nseed) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < a.total_items) a.queue[t] = t < nseed ? seeds[t] : ~0ull;
  if (t < a.total_tiles) {
    a.dep[t] = 0;
    if (a.mp_split) a.mp[t] = 0;  // part tasks finished
  }
  if (t == 0) { a.ctrl[0] = 0; a.ctrl[1] = nseed; a.ctrl[2] = 0; }
}

// max (MAXP) or min over the 8 lanes 8p..8p+7 of a group, valid on lane
// 8p: two quad permutations and a row shift by 4, as DPP operands (no LDS
// round trip).  Groups are wholly active or wholly inactive.
template <bool MAXP>
__device__ __forceinline__ int group8_reduce(int v) {
  const int a = __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xf, 0xf, false);  // quad_perm [1,0,3,2]
  v = MAXP ? imax(v, a) : imin(v, a);
  const int b = __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xf, 0xf, false);  // quad_perm [2,3,0,1]
  v = MAXP ? imax(v, b) : imin(v, b);
  const int c = __builtin_amdgcn_update_dpp(v, v, 0x104, 0xf, 0xf, false);  // row_shl:4
  return MAXP ? imax(v, c) : imin(v, c);
}

__device__ __forceinline__ int half_wave_min(int v) { return dpp_half_min(v); }

// ============================================================== Nussinov
// s(i,j) = max{0 [i==j]; s(i+1,j); s(i,j-1); s(i+1,j-1)+1 [pair(i,j-1), i+2+m<=j];
//             max_{i<k<j} s(i,k)+s(k,j)}      (bench2/nu_kern.h:23-45)
__device__ __forceinline__ bool nu_pair(int x, int y) {
  // au ua gu ug cg gc (Nussinov.scala:50-54)
  const int b = bpc(x, y);
  return b != 7;
}

struct GNu {
  static constexpr int WAVES = 4;  // waves per SIMD the register budget allows
  static constexpr int W = T + 4;  // row stride: rows start 16-byte aligned
  template <class S> __device__ static void init(const TriArgs&, S&) {}
  struct alignas(16) Smem {
    union {
      int stage[NWAVE * MP_STAGE];
      struct {
        // S: origin (IT, JT-4), so cell (IT+r, JT+c) is S[r][c+4] and a run
        // of 4 columns starting at a multiple of 4 is one 128-bit read; row
        // T and column 3 hold the frontier (row IT+32, column JT-1)
        int S[T + 1][W];
        int ST[T][W];  // the current tile transposed: ST[c][r] = S[r][c+4]
        int L[T][W];   // tile (I,I) [r][k]
        int RT[T][W];  // tile (J,J) transposed [c][k]
        int PB[T][T + 1];  // pair rule: 1 if (i, j-1) pairs and the loop is long enough, else NEG
      } w;
    } u;
    int X[T][T + 1];
  };

  struct MP {
    static constexpr bool MAXP = true, CHAIN = false, THR = false;
    static constexpr int NEUTRAL = NEG, NTv = NT_NU, LT = 0, RT = 0;
  };

  __device__ static void tile(const TriArgs& a, const TriInst& in, int I, int J, Smem& sm, int part = 0, int P = 1) {
    const int tid = threadIdx.x, IT = I * T, JT = J * T, n = in.n;
    const uint8_t* s = a.codes + in.seq_off;
    for (int e = tid; e < T * (T + 1); e += NTHR) (&sm.X[0][0])[e] = NEG;
    middle_X<MP, false>(a, in, I, J, sm.u.stage, sm.X, part, P);
    // Stage the frontier (row IT+32 and column JT-1; the current tile's
    // cells start as NEG), tiles (I,I) and (J,J), and the pair rule of
    // every cell.  The code and tile reads are issued before their LDS
    // writes, so they cost one memory latency.
    constexpr int NC = T * T / NTHR;
    uint8_t si[NC], sj[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const int e = tid + q * NTHR, i = IT + e / T, j = JT + e % T;
      const bool pl = i + 2 + a.min_loop <= j && j <= n;  // the pair rule's bounds
      si[q] = pl ? s[i] : 0;
      sj[q] = pl ? s[j - 1] : 0;
    }
    int4 l4 = make_int4(NEG, NEG, NEG, NEG), r4 = l4;
    if (I < J) {
      l4 = ((const int4*)rec_ptr(a, in, NT_NU, I, I))[tid];
      r4 = ((const int4*)rec_ptr(a, in, NT_NU, J, J))[tid];
    }
    for (int e = tid; e < (T + 1) * (T + 1); e += NTHR) {
      const int rr = e / (T + 1), cc = e % (T + 1) + 3;
      const int i = IT + rr, j = JT - 4 + cc;
      int v = NEG;
      const bool cur = rr < T && cc >= 4;
      if (!cur && i <= j && j <= n && j >= 0) v = rec_ptr(a, in, NT_NU, i / T, j / T)[(i % T) * T + (j % T)];
      sm.u.w.S[rr][cc] = v;
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const int e = tid + q * NTHR, i = IT + e / T, j = JT + e % T;
      sm.u.w.ST[e / T][e % T] = NEG;
      sm.u.w.PB[e / T][e % T] = (i + 2 + a.min_loop <= j && j <= n && nu_pair(si[q], sj[q])) ? 1 : NEG;
    }
    if (I < J) {
      const int r = (tid * 4) / T, c = (tid * 4) % T;
      *(int4*)&sm.u.w.L[r][c] = l4;
      sm.u.w.RT[c][r] = r4.x;
      sm.u.w.RT[c + 1][r] = r4.y;
      sm.u.w.RT[c + 2][r] = r4.z;
      sm.u.w.RT[c + 3][r] = r4.w;
    }
    __syncthreads();
    const int dlo = I == J ? 0 : -(T - 1);
    for (int ld = dlo; ld <= T - 1; ++ld) {
      const int rmin = ld < 0 ? -ld : 0, ncell = T - (ld < 0 ? -ld : ld);
      // split points inside blocks I and J, 8 lanes per cell (a group of 8
      // in one wave), lane `part` taking k = 4 part .. 4 part + 3 of each
      // block: its operands are four 128-bit reads (rows of L, S and the
      // transposed ST, RT), issued with the cell's own terms before any is
      // used; out-of-range candidates are masked; then a DPP max over the group
      if (tid < ncell * 8) {
        const int p = tid >> 3, part = tid & 7, r = rmin + p, c = r + ld, i = IT + r, j = JT + c, k0 = 4 * part;
        // (i+1,j), (i,j-1), (i+1,j-1); cells below the diagonal hold NEG
        const int xv = sm.X[r][c], pb = sm.u.w.PB[r][c];
        const int dn = sm.u.w.S[r + 1][c + 4], lf = sm.u.w.S[r][c + 3], dg = sm.u.w.S[r + 1][c + 3];
        const int4 sr = *(const int4*)&sm.u.w.S[r][k0 + 4];  // s(i, JT+k0+m)
        const int4 sc = *(const int4*)&sm.u.w.ST[c][k0];     // s(IT+k0+m, j)
        int best = NEG;
        if (I == J) {  // k = IT + kk, i < k < j
          const int x[4] = {sr.x + sc.x, sr.y + sc.y, sr.z + sc.z, sr.w + sc.w};
#pragma unroll
          for (int m = 0; m < 4; ++m) best = imax(best, k0 + m > r && k0 + m < c ? x[m] : NEG);
        } else {  // k = IT + kk > i with s(i,k) from L, and k = JT + kk < j with s(k,j) from RT
          const int4 lv = *(const int4*)&sm.u.w.L[r][k0];
          const int4 rv = *(const int4*)&sm.u.w.RT[c][k0];
          const int x1[4] = {lv.x + sc.x, lv.y + sc.y, lv.z + sc.z, lv.w + sc.w};
          const int x2[4] = {sr.x + rv.x, sr.y + rv.y, sr.z + rv.z, sr.w + rv.w};
#pragma unroll
          for (int m = 0; m < 4; ++m)
            best = imax(best, imax(k0 + m > r ? x1[m] : NEG, k0 + m < c ? x2[m] : NEG));
        }
        if (!(i < j && j <= n)) best = NEG;
        best = group8_reduce<true>(best);
        // the cell's value on the group's first lane (the cells of this
        // diagonal read only earlier diagonals).  For i < j, s(i+1,j) >= 0
        // is a candidate, so NEG-based sums never win; for i == j the
        // neighbours are below the diagonal (NEG)
        if (part == 0) {
          int v = NEG;
          if (i <= j && j <= n) {
            v = imax(imax(xv, best), imax(dn, lf));
            if (i == j) v = imax(v, 0);
            v = imax(v, dg + pb);
          }
          sm.u.w.S[r][c + 4] = v;
          sm.u.w.ST[c][r] = v;
        }
      }
      __syncthreads();
    }
    int* g = rec_ptr(a, in, NT_NU, I, J);
    for (int e = tid; e < T * T; e += NTHR) g[e] = sm.u.w.S[e / T][e % T + 4];
  }
};

// ============================================================ matrix chain
// M(i,j) = min_{i<k<j} M(i,k)+M(k,j)+rows[i]*rows[k]*cols[j-1]  (int32 wrap),
// M(i,i+1) = 0   (Benchmarks.scala:18-27)
struct GMc {
  static constexpr int WAVES = 4;  // waves per SIMD the register budget allows
  template <class S> __device__ static void init(const TriArgs&, S&) {}
  struct Smem {
    union {
      int stage[NWAVE * MP_STAGE];
      struct {
        int S[T][T + 1];      // current tile
        int L[T][T + 1];      // tile (I,I)
        int R[T][T + 1];      // tile (J,J)
        int rows[2 * T];      // rows[IT..IT+T) , rows[JT..JT+T)
        uint32_t AL[T][T + 1];  // rows[i] * cols[j-1] per cell
      } w;
    } u;
    int X[T][T + 1];
  };

  struct MP {
    static constexpr bool MAXP = false, CHAIN = true, THR = false;
    static constexpr int NEUTRAL = 0x7fffffff, NTv = NT_MC, LT = 0, RT = 0;
  };

  __device__ static void tile(const TriArgs& a, const TriInst& in, int I, int J, Smem& sm, int part = 0, int P = 1) {
    const int tid = threadIdx.x, IT = I * T, JT = J * T, n = in.n;
    const int32_t* rows = a.dims + in.seq_off;
    const int32_t* cols = rows + n;
    for (int e = tid; e < T * (T + 1); e += NTHR) (&sm.X[0][0])[e] = 0x7fffffff;
    middle_X<MP, false>(a, in, I, J, sm.u.stage, sm.X, part, P);
    for (int e = tid; e < T * (T + 1); e += NTHR) (&sm.u.w.S[0][0])[e] = 0;
    // global reads first, then the LDS writes (one memory latency)
    constexpr int NC = T * T / NTHR;
    uint32_t ri[NC], cj[NC];
    int lv[NC], rv[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const int e = tid + q * NTHR, i = IT + e / T, j = JT + e % T;
      const bool live = i + 1 < j && j <= n;
      ri[q] = live ? (uint32_t)rows[i] : 0u;
      cj[q] = live ? (uint32_t)cols[j - 1] : 0u;
      lv[q] = I < J ? rec_ptr(a, in, NT_MC, I, I)[e] : 0;
      rv[q] = I < J ? rec_ptr(a, in, NT_MC, J, J)[e] : 0;
    }
    if (I < J) {
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        const int e = tid + q * NTHR;
        sm.u.w.L[e / T][e % T] = lv[q];
        sm.u.w.R[e / T][e % T] = rv[q];
      }
    }
    int rk = 0;
    if (tid < 2 * T) {
      const int k = tid < T ? IT + tid : JT + tid - T;
      rk = k < n ? rows[k] : 0;
    }
    if (tid < 2 * T) sm.u.w.rows[tid] = rk;
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const int e = tid + q * NTHR;
      sm.u.w.AL[e / T][e % T] = ri[q] * cj[q];
    }
    __syncthreads();
    const int dlo = I == J ? 0 : -(T - 1);
    for (int ld = dlo; ld <= T - 1; ++ld) {
      const int rmin = ld < 0 ? -ld : 0, ncell = T - (ld < 0 ? -ld : ld);
      // split points inside blocks I and J: 8 lanes per cell, strided
      // shares with independent loads, DPP min over the group (int32 wrap:
      // the minimum of the wrapped candidates, as the sequential loop keeps)
      if (tid < ncell * 8) {
        const int p = tid >> 3, part = tid & 7, r = rmin + p, c = r + ld, i = IT + r, j = JT + c;
        const int xv = sm.X[r][c];
        const uint32_t alpha = sm.u.w.AL[r][c];
        const bool live = i + 1 < j && j <= n;
        // kk = part + 8m (m < 4) per range, fixed trip counts, masked after the loads
        int best = 0x7fffffff;
        if (I == J) {
#pragma unroll
          for (int m = 0; m < 4; ++m) {  // k = IT + kk, i < k < j
            const int kk = part + 8 * m;
            const int x = (int)((uint32_t)sm.u.w.S[r][kk] + (uint32_t)sm.u.w.S[kk][c] + alpha * (uint32_t)sm.u.w.rows[kk]);
            best = imin(best, kk > r && kk < c ? x : 0x7fffffff);
          }
        } else {
#pragma unroll
          for (int m = 0; m < 4; ++m) {  // k = IT + kk > i, and k = JT + kk < j
            const int kk = part + 8 * m;
            const int x1 = (int)((uint32_t)sm.u.w.L[r][kk] + (uint32_t)sm.u.w.S[kk][c] + alpha * (uint32_t)sm.u.w.rows[kk]);
            const int x2 =
                (int)((uint32_t)sm.u.w.S[r][kk] + (uint32_t)sm.u.w.R[kk][c] + alpha * (uint32_t)sm.u.w.rows[T + kk]);
            best = imin(best, imin(kk > r ? x1 : 0x7fffffff, kk < c ? x2 : 0x7fffffff));
          }
        }
        if (!live) best = 0x7fffffff;
        best = group8_reduce<false>(best);
        // the cell's value on the group's first lane (this diagonal's cells
        // read only earlier diagonals): M(i,i+1) = 0, else the minimum
        if (part == 0) sm.u.w.S[r][c] = live ? imin(xv, best) : 0;
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
// Energy tables read by the loop phase of the Zuker fill, staged once per
// workgroup in LDS (int16, the width of the reference's paramT,
// vienna.h:24-67); the tables used only by the sequence terms are read from
// the device copy of the parameters.
struct ZuTab {
  int16_t stack[8][8];
  int16_t mmI[8][5][5], mm1nI[8][5][5], mm23I[8][5][5];
  int16_t bulge[32];
  int32_t term_au, il23;  // terminal AU penalty; 2x3 loop energy il(5) + ninio
};

// Per-cell terms that depend on the sequence only, precomputed one wavefront
// step ahead by the waves that are idle during the cell-finalize phase.
struct ZuPre {
  int16_t valid;  // cl(i,j) exists: j-i >= 7 and both stacked pairs canonical (fun6)
  int16_t sr;     // sr_energy(i,j-1)
  int16_t r1;     // rule 1: hairpin(i+1,j-2) + sr
  int16_t r5c;    // rule 5 without ml: ML_closing + ML_intern + termau + ml_mismatch + sr
  int16_t uoff;   // U - cl = ext_mismatch(i,j-1) + termau(i,j-1) + ML_intern
  int16_t boff;   // B - cl = interior mismatch of (i,j) as an inner pair (kNoB if unused)
  int16_t outI;   // generic interior-loop outer mismatch + sr
  int16_t in0;    // interior mismatch of (i+1,j-1): recovers cl(i+1,j-1) from B (rule 0)
  // 1x1, 1x2, 2x1, 2x2 loops: il_energy (int11/int21/int22) minus the inner
  // pair's interior mismatch, so that the candidate is B(l,k) + f + sr
  int16_t f11, f12, f21, f22;
};
constexpr int kNoB = -32768;

// A tile's view of the sequence: row side [IT-1, IT+66), column side [JT-34, JT+34).
struct ZuView {
  const ZuTab& E;
  const DevRna* P;
  const int32_t* hl;
  const uint8_t* rs;
  const uint8_t* cs;
  int IT1, JT34, n;
  const int16_t* sp;  // special-hairpin energies of this instance, [cls * spn + position]
  int64_t spn;
  __device__ int R(int x) const { return rs[x - IT1]; }
  // energy of the special tri- (cls 0), tetra- (1) or hexaloop (2) closed by
  // the pair starting at position x, SPEC_NONE if there is none
  __device__ int spec(int x, int cls) const { return sp[cls * spn + x]; }
  __device__ int C(int x) const { return cs[x - JT34]; }
  __device__ int tau(int x, int y) const {  // termau_energy on bases (librna_impl.h:152)
    return ((x == 3 && y == 2) || (x == 2 && y == 3)) ? 0 : E.term_au;
  }
  __device__ bool stackpair(int i, int j) const {  // fun6 (zu.c:35)
    return i + 4 <= j && bpc(R(i), C(j - 1)) != 7 && bpc(R(i + 1), C(j - 2)) != 7;
  }
  // interior mismatch of the pair closing cell (l,k) (pair l..k-1), seen from outside
  __device__ int inner(int l, int k) const { return E.mmI[bpc(C(k - 1), R(l))][C(k)][R(l - 1)]; }
  __device__ int sr(int i, int j) const { return E.stack[bpc(R(i), C(j - 1))][bpc(C(j - 2), R(i + 1))]; }
  // Sequence-only terms of cell (i,j).  The device-memory table reads are
  // independent of each other, so they are all issued before any is used.
  // Returns whether cl(i,j) exists (o->valid).
  __device__ bool pre(int i, int j, ZuPre* o) const {
    const int d = j - i;
    o->valid = 0;
    if (j > n || d < 7 || !stackpair(i, j)) return false;
    o->valid = 1;
    const int Ri = R(i), Ri1 = R(i + 1), Ri2 = R(i + 2), Ri3 = R(i + 3), Ri4 = R(i + 4);
    const int Cj1 = C(j - 1), Cj2 = C(j - 2), Cj3 = C(j - 3), Cj4 = C(j - 4), Cj5 = C(j - 5);
    const int ocb = bpc(Ri1, Cj2), bo = bpc(Ri, Cj1);
    // hl_energy(i+1, j-2) generic part (librna_impl.h:181-193), loop size d-4
    const int size = d - 4;
    const int hent = hl[size];  // hairpin[size] up to 30, extrapolated beyond
    const int hmm = P->mmH[ocb][Ri2][Cj3];
    // ml_mismatch of the closing pair (rule 5), ext_mismatch(i, j-1) (rule 10/U)
    const int mlm = P->mmM[bpc(Cj2, Ri1)][Cj3][Ri2];
    const int* ep = (i > 0 && j < n) ? &P->mmExt[bo][R(i - 1)][C(j)] : (i > 0 ? &P->d5[bo][R(i - 1)] : &P->d3[bo][C(j)]);
    const int ex = (i > 0 || j < n) ? *ep : 0;
    // 1x1, 1x2, 2x1, 2x2 loop energies (il11/12/21/22_energy, librna_impl.h:222-225):
    // read for every cell with the other loads (the indices are base codes
    // and pair types, always inside the tables), used when d >= 12
    const int e11 = P->int11[ocb][bpc(Cj4, Ri3)][Ri2][Cj3];
    const int e12 = P->int21[ocb][bpc(Cj5, Ri3)][Ri2][Cj4][Cj3];
    const int e21 = P->int21[bpc(Cj4, Ri4)][ocb][Cj3][Ri2][Ri3];
    const int e22 = P->int22[ocb][bpc(Cj5, Ri4)][Ri2][Ri3][Cj4][Cj3];
    int hp = size == 3 ? hent + tau(Ri1, Cj2) : hent + hmm;
    const int cls = size == 3 ? 0 : size == 4 ? 1 : size == 6 ? 2 : -1;
    if (cls >= 0) {  // special tri-/tetra-/hexaloops, matched once per position by tri_spec
      const int sv = spec(i + 1, cls);
      if (sv != SPEC_NONE) hp = sv;
    }
    const int s_ = E.stack[bo][bpc(Cj2, Ri1)];
    o->sr = s_;
    o->r1 = hp + s_;
    o->r5c = P->ml_closing + P->ml_intern + tau(Ri1, Cj2) + mlm + s_;
    o->uoff = ex + tau(Ri, Cj1) + P->ml_intern;
    o->boff = (i >= 1 && j <= n - 1) ? inner(i, j) : kNoB;
    o->outI = E.mmI[ocb][Ri2][Cj3] + s_;
    o->in0 = d >= 9 ? inner(i + 1, j - 1) : 0;
    if (d >= 12) {
      o->f11 = (int16_t)(e11 - inner(i + 3, j - 3));
      o->f12 = (int16_t)(e12 - inner(i + 3, j - 4));
      o->f21 = (int16_t)(e21 - inner(i + 4, j - 3));
      o->f22 = (int16_t)(e22 - inner(i + 4, j - 4));
    }
    return true;
  }
};

// energy tables of the loop phases, loaded once per workgroup
__device__ void load_tab(const TriArgs& a, ZuTab& E) {
  const DevRna* P = (const DevRna*)a.rna;
  for (int e = threadIdx.x; e < 64; e += NTHR) (&E.stack[0][0])[e] = (int16_t)(&P->stack[0][0])[e];
  for (int e = threadIdx.x; e < 200; e += NTHR) {
    (&E.mmI[0][0][0])[e] = (int16_t)(&P->mmI[0][0][0])[e];
    (&E.mm1nI[0][0][0])[e] = (int16_t)(&P->mm1nI[0][0][0])[e];
    (&E.mm23I[0][0][0])[e] = (int16_t)(&P->mm23I[0][0][0])[e];
  }
  for (int e = threadIdx.x; e < 31; e += NTHR) E.bulge[e] = (int16_t)P->bulge[e];
  if (threadIdx.x == 0) {
    E.term_au = P->term_au;
    E.il23 = P->il5 + P->ninio0;
  }
}

struct GZu {
  static constexpr int WAVES = 3;  // waves per SIMD the register budget allows
  static constexpr int BR = 64, BC = 64;   // B region: rows IT..IT+63, cols JT-32..JT+31
  static constexpr int BS = BC + 1;        // odd row stride: lanes walking a column hit distinct banks
  // The interior-loop window of a cell on diagonal d reads diagonals <= d-6,
  // its bulges <= d-5 and its split points inside blocks I, J <= d-7; so the
  // heavy terms of G consecutive diagonals are computed in one wide phase, and
  // only the O(1) cell finalize (rules 0, 1, 5, 6, 8, 10 and the prefix mins)
  // runs diagonal by diagonal.
  static constexpr int G = 5;
  struct Smem {
    union {
      int stage[NWAVE * MP_STAGE];
      struct {
        int B[BR][BS];         // cl + inner mismatch, origin (IT, JT-32)
        int U[T][T + 1];       // origin (IT, JT-1)
        int ML[T + 2][T + 2];  // origin (IT, JT-2)
        int ML1[T + 1][T + 1]; // origin (IT, JT)
        int R11[T][T + 1];     // origin (IT, JT-1)
        int RM1[T][T + 1];     // ml1 tile (J,J)
      } w;
    } u;
    int X[T][T + 1];
    ZuTab E;
    ZuPre pre[2][G][T];  // double buffer: light steps of one band, sequence terms of the next
    uint8_t rs[72], cs[72];
    int acc[G][T];
    int16_t list[G * T];
    int nlist[2];  // by band parity: the next band's list is counted while this one's is read
  };

  __device__ static void init(const TriArgs& a, Smem& sm) { load_tab(a, sm.E); }

  // Sequence-only terms of the cells of band [ld0, ld0+G) into pre[buf], and
  // the list of cells whose cl can close bulges / interior loops (the cells
  // with cl and j - i >= 12), in any order: each wave appends its cells with
  // one LDS atomic on nlist[buf], which must be 0 beforehand.  Called by
  // whole waves, t in [0, 192) (cell t of the band's G*T <= 160).
  __device__ static void prep_band(const ZuView& z, int IT, int JT, int ld0, Smem& sm, int t, int buf) {
    bool listed = false;
    const int q = t / T, p = t % T, ld = ld0 + q;
    if (t < G * T) {
      const int rmin = ld < 0 ? -ld : 0, ncell = T - (ld < 0 ? -ld : ld);
      if (ld <= T - 1 && p < ncell) {
        const int r = rmin + p;
        listed = z.pre(IT + r, JT + r + ld, &sm.pre[buf][q][p]) && ld + JT - IT >= 12;
      }
    }
    const unsigned long long m = __ballot(listed);
    if (m) {
      int base = 0;
      if ((threadIdx.x & 63) == 0) base = atomicAdd(&sm.nlist[buf], __popcll(m));
      base = __builtin_amdgcn_readfirstlane(base);
      if (listed) sm.list[base + __popcll(m & ((1ull << (threadIdx.x & 63)) - 1))] = (int16_t)(q * T + p);
    }
  }

  static constexpr int kMask = 1 << 28;  // masked loop energy: candidates through it stay >= INV_THR

  // Rules 2, 3 and 4 of listed cell (i,j) (tile coordinates r, c): lane ap of
  // the cell's 32 lanes (a half wave) evaluates its share, the result is the
  // minimum over the 32 lanes, returned on lanes ap >= 16.  Lane ap < 30 (left size a = ap+1): the generic interior
  // loops (a, b = 2..30), the loops (a,1) and (1,a) for a >= 2, and the left
  // and right bulges of size a.  Lane 30: (1,1) and (2,2); lane 31: (2,3) and
  // (3,2).  B(l,k) = cl(l,k) + interior mismatch of (l,k), so
  // cl(l,k) = B(l,k) - inner(l,k) (librna_impl.h:199-265).
  template <class S>
  __device__ static int window_lane(const ZuView& z, S& sm, const ZuPre& pc, int ap, int i, int j, int r, int c,
                                    const int* fg, int f_a1, int f_1a, int amax = 30) {
    const ZuTab& E = sm.E;
    // outer pair (i+1, j-2) of the loops and its mismatch bases
    const int Ri1 = z.R(i + 1), Ri2 = z.R(i + 2), Cj2 = z.C(j - 2), Cj3 = z.C(j - 3);
    const int ocb = bpc(Ri1, Cj2);
    int best = INV;
    if (ap < amax) {
      const int aa = ap + 1;
      const int* Brow = &sm.u.w.B[r + 2 + aa][c + 30];  // Brow[-b] = B(i+2+aa, j-2-b)
      int g = INV;
      // all 29 loads issued before the adds: the scheduler would otherwise
      // reuse two registers and wait for every load in turn
      int v[29];
#pragma unroll
      for (int b = 2; b <= 30; ++b) v[b - 2] = Brow[-b];
      __builtin_amdgcn_sched_group_barrier(0x100, 15, 0);  // LDS reads
      __builtin_amdgcn_sched_group_barrier(0x002, 58, 0);  // VALU
#pragma unroll
      for (int b = 2; b <= 30; ++b) g = imin(g, v[b - 2] + fg[b - 2]);
      best = g + pc.outI;
      const int Rl = z.R(i + 2 + aa), Rl1 = z.R(i + 1 + aa);
      const int out1n = E.mm1nI[ocb][Ri2][Cj3] + pc.sr, tauo = z.tau(Ri1, Cj2);
      if (aa >= 2) {
        // (aa, 1): inner pair (i+2+aa, j-3); 2x1 from int21, else 1xn
        const int Bv = Brow[-1], icb = bpc(z.C(j - 4), Rl);
        const int cl = Bv - E.mmI[icb][Cj3][Rl1];
        const int cand = aa == 2 ? Bv + pc.f21 + pc.sr : cl + f_a1 + E.mm1nI[icb][Cj3][Rl1] + out1n;
        if (Bv < INV_THR) best = imin(best, cand);
        // (1, aa): inner pair (i+3, j-2-aa); 1x2 from int21, else 1xn
        const int k = j - 2 - aa, Bw = sm.u.w.B[r + 3][c + 30 - aa];
        const int Ck = z.C(k), icb2 = bpc(z.C(k - 1), z.R(i + 3));
        const int cl2 = Bw - E.mmI[icb2][Ck][Ri2];
        const int cand2 = aa == 2 ? Bw + pc.f12 + pc.sr : cl2 + f_1a + E.mm1nI[icb2][Ck][Ri2] + out1n;
        if (Bw < INV_THR) best = imin(best, cand2);
      }
      {
        // left bulge of size aa: inner pair (i+2+aa, j-2)
        const int Bv = Brow[0], icb = bpc(Cj3, Rl);
        const int cl = Bv - E.mmI[icb][Cj2][Rl1];
        const int e = aa == 1 ? E.bulge[1] + E.stack[ocb][icb] : E.bulge[aa] + tauo + z.tau(Cj3, Rl);
        if (Bv < INV_THR) best = imin(best, cl + e + pc.sr);
      }
      {
        // right bulge of size aa: inner pair (i+2, j-2-aa)
        const int k = j - 2 - aa, Bv = sm.u.w.B[r + 2][c + 30 - aa];
        const int Ck1 = z.C(k - 1), icb = bpc(Ck1, Ri2);
        const int cl = Bv - E.mmI[icb][z.C(k)][Ri1];
        const int e = aa == 1 ? E.bulge[1] + E.stack[ocb][icb] : E.bulge[aa] + tauo + z.tau(Ck1, Ri2);
        if (Bv < INV_THR) best = imin(best, cl + e + pc.sr);
      }
    } else if (ap < 30) {
      // rows beyond amax: nothing
    } else if (ap == 30) {
      // 1x1 and 2x2 loops (int11, int22)
      const int B11 = sm.u.w.B[r + 3][c + 29], B22 = sm.u.w.B[r + 4][c + 28];
      if (B11 < INV_THR) best = imin(best, B11 + pc.f11 + pc.sr);
      if (B22 < INV_THR) best = imin(best, B22 + pc.f22 + pc.sr);
    } else {
      // 2x3 and 3x2 loops
      const int B23 = sm.u.w.B[r + 4][c + 27], B32 = sm.u.w.B[r + 5][c + 28];
      const int out23 = E.il23 + E.mm23I[ocb][Ri2][Cj3] + pc.sr;
      if (B23 < INV_THR) {
        const int l = i + 4, k = j - 5, icb = bpc(z.C(k - 1), z.R(l)), Ck = z.C(k), Rl1 = z.R(l - 1);
        best = imin(best, B23 - E.mmI[icb][Ck][Rl1] + E.mm23I[icb][Ck][Rl1] + out23);
      }
      if (B32 < INV_THR) {
        const int l = i + 5, k = j - 4, icb = bpc(z.C(k - 1), z.R(l)), Ck = z.C(k), Rl1 = z.R(l - 1);
        best = imin(best, B32 - E.mmI[icb][Ck][Rl1] + E.mm23I[icb][Ck][Rl1] + out23);
      }
    }
    return half_wave_min(best);  // valid on lanes ap >= 16
  }

  struct MP {
    static constexpr bool MAXP = false, CHAIN = false, THR = true;
    static constexpr int NEUTRAL = INV, NTv = NT_ZU, LT = ZT_U, RT = ZT_ML1;
  };

  __device__ static void tile(const TriArgs& a, const TriInst& in, int I, int J, Smem& sm, int part = 0, int nparts = 1) {
    const int tid = threadIdx.x, IT = I * T, JT = J * T, n = in.n, nb = in.nb;
    const uint8_t* seq = a.codes + in.seq_off;
    const DevRna* P = (const DevRna*)a.rna;
    const ZuView z{sm.E, P, a.hl_ent, sm.rs, sm.cs, IT - 1, JT - 34, n, a.spec + in.seq_off, a.n_codes};
    for (int e = tid; e < T * (T + 1); e += NTHR) (&sm.X[0][0])[e] = INV;
    for (int e = tid; e < G * T; e += NTHR) (&sm.acc[0][0])[e] = INV;
    if (tid == 0) sm.nlist[0] = 0;
    if (tid < 72) {
      const int x = IT - 1 + tid, y = JT - 34 + tid;
      sm.rs[tid] = (x >= 0 && x < n) ? seq[x] : 0;
      sm.cs[tid] = (y >= 0 && y < n) ? seq[y] : 0;
    }
    middle_X<MP, false>(a, in, I, J, sm.u.stage, sm.X, part, nparts);
    // ---- stage the frontier (values of finished neighbour tiles) ----
    auto fetch = [&](int tab, int i, int j) -> int {
      if (i < 0 || j < 0 || i > j || j > n) return INV;
      const int I2 = i / T, J2 = j / T;
      if (I2 >= nb || J2 >= nb) return INV;
      return rec_ptr(a, in, NT_ZU, I2, J2)[tab * T * T + (i % T) * T + (j % T)];
    };
    // Every global read of the frontier is issued before the first LDS
    // write, so the workgroup waits for one memory latency, not one per loop.
    // B quadrants: tiles (I,J-1) | current tile / (I+1,J-1) | (I+1,J); absent tiles are INV
    const int* qb[3] = {J - 1 >= I ? rec_ptr(a, in, NT_ZU, I, J - 1) + ZT_B * T * T : nullptr,
                        I + 1 <= J - 1 ? rec_ptr(a, in, NT_ZU, I + 1, J - 1) + ZT_B * T * T : nullptr,
                        I + 1 <= J ? rec_ptr(a, in, NT_ZU, I + 1, J) + ZT_B * T * T : nullptr};
    int4 qv[3];
#pragma unroll
    for (int quad = 0; quad < 3; ++quad)
      qv[quad] = qb[quad] ? *(const int4*)(qb[quad] + tid * 4) : make_int4(INV, INV, INV, INV);
    const int4 rm = I < J ? *(const int4*)(rec_ptr(a, in, NT_ZU, J, J) + ZT_ML1 * T * T + tid * 4)
                          : make_int4(INV, INV, INV, INV);  // ml1 tile (J,J)
    // one boundary cell per thread: [0,32) U and R11 at column JT-1; [32,100)
    // ML rows IT+32, IT+33; [100,164) ML columns JT-2, JT-1; [164,196) ML1 row IT+32
    int bv0 = INV, bv1 = INV;
    if (tid < 32) {
      bv0 = fetch(ZT_U, IT + tid, JT - 1);
      bv1 = fetch(ZT_R11, IT + tid, JT - 1);
    } else if (tid < 100) {
      const int t = tid - 32;
      bv0 = fetch(ZT_ML, IT + T + t / (T + 2), JT - 2 + t % (T + 2));
    } else if (tid < 164) {
      const int t = tid - 100;
      bv0 = fetch(ZT_ML, IT + (t >> 1), JT - 2 + (t & 1));
    } else if (tid < 196) {
      bv0 = fetch(ZT_ML1, IT + T, JT + tid - 164);
    }
    // the current tile's cells start empty
    for (int e = tid; e < T * T; e += NTHR) sm.u.w.B[e / T][T + e % T] = INV;
    for (int e = tid; e < T * (T + 1); e += NTHR) {
      const int rr = e / (T + 1), cc = e % (T + 1);
      if (cc != 0) sm.u.w.U[rr][cc] = sm.u.w.R11[rr][cc] = INV;
    }
    for (int e = tid; e < (T + 2) * (T + 2); e += NTHR) {
      const int rr = e / (T + 2), cc = e % (T + 2);
      if (rr < T && cc >= 2) sm.u.w.ML[rr][cc] = INV;
    }
    for (int e = tid; e < T * T; e += NTHR) sm.u.w.ML1[e / T][e % T] = INV;
    // then the values read above
#pragma unroll
    for (int quad = 0; quad < 3; ++quad) {
      const int x = tid * 4, rr = x / T + (quad ? T : 0), cc = x % T + (quad == 2 ? T : 0);
      int* dst = &sm.u.w.B[rr][cc];
      dst[0] = qv[quad].x;
      dst[1] = qv[quad].y;
      dst[2] = qv[quad].z;
      dst[3] = qv[quad].w;
    }
    if (I < J) {
      int* dst = &sm.u.w.RM1[(tid * 4) / T][(tid * 4) % T];
      dst[0] = rm.x;
      dst[1] = rm.y;
      dst[2] = rm.z;
      dst[3] = rm.w;
    }
    if (tid < 32) {
      sm.u.w.U[tid][0] = bv0;
      sm.u.w.R11[tid][0] = bv1;
    } else if (tid < 100) {
      const int t = tid - 32;
      sm.u.w.ML[T + t / (T + 2)][t % (T + 2)] = bv0;
    } else if (tid < 164) {
      const int t = tid - 100;
      sm.u.w.ML[t >> 1][t & 1] = bv0;
    } else if (tid < 196) {
      sm.u.w.ML1[T][tid - 164] = bv0;
    }
    const int* gLU = rec_ptr(a, in, NT_ZU, I, I) + ZT_U * T * T;  // U tile (I,I), read in place
    const int dlo = I == J ? 0 : -(T - 1);
    __syncthreads();  // sequence segments visible to prep_band
    if (tid >= 64) prep_band(z, IT, JT, dlo, sm, tid - 64, 0);
    __syncthreads();
    int32_t* gcl = rec_ptr(a, in, NT_ZU, I, J) + ZT_CL * T * T;
    // interior-loop lane of this thread: left size a = ap + 1 (ap < 30); its
    // row of generic loop energies F[a][2..30] in registers, the special
    // positions (all of a == 1, (2,2), (2,3), (3,2)) masked with kMask
    const int ap = tid & 31, aa = ap + 1;
    int fg[29];
#pragma unroll
    for (int b = 2; b <= 30; ++b) {
      const bool special = ap >= 30 || aa == 1 || (aa == 2 && b <= 3) || (aa == 3 && b == 2);
      fg[b - 2] = special ? kMask : P->F[ap < 30 ? aa : 1][b];
    }
    const int f_a1 = P->F[ap < 30 ? aa : 1][1], f_1a = P->F[1][ap < 30 ? aa : 1];
    for (int ld0 = dlo, buf = 0; ld0 <= T - 1; ld0 += G, buf ^= 1) {
      // ---------------- heavy phase: band [ld0, ld0+G) ----------------
      // Lanes [0, wl): 32 per listed cell (its interior loops and bulges),
      // padded to whole waves; lanes [wl, wl + 2*ncell): two per cell of the
      // band (split points of X inside block I, inside block J).  Every wave
      // runs one kind of lane with fixed trip counts: the operands outside a
      // rule's index range are INV in the tables, so no bounds are tested.
      const int qn = imin(G, T - ld0);
      int cpre[G + 1];
      cpre[0] = 0;
      for (int q = 0; q < G; ++q) cpre[q + 1] = cpre[q] + (q < qn ? T - abs(ld0 + q) : 0);
      const int nv = sm.nlist[buf], wl = ((nv + 1) >> 1) * 64, total = wl + 2 * cpre[G];
      if (tid == 0) sm.nlist[buf ^ 1] = 0;  // counted by the next band's prep, after the barrier
      for (int w = tid; w < total; w += NTHR) {
        if (w < wl) {
          const int v = w >> 5;
          if (v >= nv) continue;
          const int qp = sm.list[v], q = qp / T, p = qp % T, ld = ld0 + q;
          const int r = (ld < 0 ? -ld : 0) + p, c = r + ld, i = IT + r, j = JT + c;
          const int best = window_lane(z, sm, sm.pre[buf][q][p], ap, i, j, r, c, fg, f_a1, f_1a);
          if (ap == 31) sm.acc[q][p] = best;
        } else {
          const int e = (w - wl) >> 1, part = w & 1;
          int q = 0;
          while (e >= cpre[q + 1]) ++q;
          const int ld = ld0 + q, r = (ld < 0 ? -ld : 0) + (e - cpre[q]), c = r + ld;
          int best = INV;
          if (part == 0) {
            if (I == J) {
              best = lds_minplus8(&sm.u.w.U[r][1], &sm.u.w.ML1[0][c], T + 1);
            } else {
              // U(I,I) row r from HBM/L2: the whole row is requested at
              // once, so the lane waits for one memory latency, not four
              const int4* g4 = (const int4*)(gLU + r * T);
              int4 u4[T / 4];
#pragma unroll
              for (int x4 = 0; x4 < T / 4; ++x4) u4[x4] = g4[x4];
#pragma unroll
              for (int x0 = 0; x0 < T; x0 += 8) {
                int bv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) bv[u] = sm.u.w.ML1[x0 + u][c];
                const int4 ua = u4[x0 / 4], ub = u4[x0 / 4 + 1];
                best = imin(best, imin(imin(ua.x + bv[0], ua.y + bv[1]), imin(ua.z + bv[2], ua.w + bv[3])));
                best = imin(best, imin(imin(ub.x + bv[4], ub.y + bv[5]), imin(ub.z + bv[6], ub.w + bv[7])));
              }
            }
          } else if (I < J) {
            best = lds_minplus8(&sm.u.w.U[r][1], &sm.u.w.RM1[0][c], T + 1);
          }
          best = imin(best, __builtin_amdgcn_update_dpp(best, best, 0xB1, 0xf, 0xf, false));  // the pair's other lane
          if (part == 0 && best < INV_THR) sm.X[r][c] = imin(sm.X[r][c], best);
        }
      }
      __syncthreads();
      // ---------------- light steps: finalize the band diagonal by diagonal ----------------
      // Wave 0 alone walks the band's diagonals (ordered by its own LDS
      // writes, no workgroup barrier); waves 1-3 meanwhile compute the
      // sequence terms and cell list of the next band into the other buffer.
      if (tid >= 64) {
        if (ld0 + G <= T - 1) prep_band(z, IT, JT, ld0 + G, sm, tid - 64, buf ^ 1);
      } else {
        for (int q = 0; q < qn; ++q) {
          const int ld = ld0 + q, rmin = ld < 0 ? -ld : 0, ncell = T - (ld < 0 ? -ld : ld);
          if (tid < ncell) {
          const int p = tid, r = rmin + p, c = r + ld, i = IT + r, j = JT + c, d = j - i;
          // every operand first (finished diagonals and this cell's own
          // terms; all addresses inside the arrays), then branch-free
          // arithmetic: one LDS round trip per diagonal step
          const ZuPre pc = sm.pre[buf][q][p];
          const int accv = sm.acc[q][p], Bi = sm.u.w.B[r + 1][c + 31], ml2 = sm.u.w.ML[r + 2][c];
          const int Xv = sm.X[r][c], mlp = sm.u.w.ML[r + 1][c + 2], m1p = sm.u.w.ML1[r + 1][c];
          const int Rp = sm.u.w.R11[r][c], Up = sm.u.w.U[r][c];
          const bool live = i <= j && j <= n, v = live && pc.valid;
          int cl = imin(accv, pc.r1);
          cl = d >= 9 && Bi < INV_THR ? imin(cl, Bi - pc.in0 + pc.sr) : cl;
          cl = d >= 18 && ml2 < INV_THR ? imin(cl, ml2 + pc.r5c) : cl;
          cl = v ? cl : INV;
          const int U = v ? cl + pc.uoff : INV;
          const int Bv = v && pc.boff != kNoB ? cl + pc.boff : INV;
          const int X = Xv >= INV_THR ? INV : Xv;
          int ml = d >= 15 ? imin(X, mlp) : X;
          int R11 = d >= 8 ? (d >= 9 ? imin(Rp, Up) : Up) : INV;
          int ml1 = imin(imin(X, U), R11);
          ml1 = d >= 8 ? imin(ml1, m1p) : ml1;
          ml = live && ml < INV_THR ? ml : INV;
          ml1 = live && ml1 < INV_THR ? ml1 : INV;
          R11 = live && R11 < INV_THR ? R11 : INV;
          sm.u.w.B[r][c + 32] = Bv;
          sm.u.w.U[r][c + 1] = U;
          sm.u.w.ML[r][c + 2] = ml;
          sm.u.w.ML1[r][c] = ml1;
          sm.u.w.R11[r][c + 1] = R11;
          gcl[r * T + c] = cl;
          sm.acc[q][p] = INV;
          }
          wave_lds_sync();
        }
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
      g[ZT_B * T * T + e] = sm.u.w.B[r][T + c];
    }
  }
};

// ============================================================== RNAfold
// RNAfold grammar with the MFE algebra (RNAFold.scala:20-64), tables per tile
// Q, QM, Qp, B:
//   Qp(i,j) [pair (i,j-1)] = min{ hl(i,j-1); sr(i,j-1) + Qp(i+1,j-1);
//                                 il(i,k,l-1,j-1) + Qp(k,l), 1<=k-i<=30, 1<=j-l<=30; QM(i+1,j-1) }
//   QM(i,j) = min_k Q(i,k) + Q(k,j)                                   [j-i >= 4]
//   Q(i,j)  = min{ QM(i,j); Q(i,j-1); Q(i+1,j); Qp(i,j) }               [j-i >= 2]
// The interior-loop window of Qp(i,j) (outer pair (i, j-1), left / right
// unpaired sizes sl, sr in [0,29]) is the loop window of the Zuker closed
// table at (i-1, j+1) with sizes a = sl, b = sr, the sl = 0 / sr = 0 rows
// being its bulges; so the same 32-lane evaluation is used with the stacked
// pair's energy set to zero and sizes capped at 29.  (0,0) is the stack rule.
struct GRf {
  static constexpr int WAVES = 3;  // waves per SIMD the register budget allows
  static constexpr int BR = 64, BS = 65;
  // window candidates of a cell on diagonal d read diagonals <= d-3 (except
  // (0,0), the stack), the split points of QM <= d-3: bands of 3 diagonals
  static constexpr int G = 3;
  struct Smem {
    union {
      int stage[NWAVE * MP_STAGE];
      struct {
        int B[BR][BS];          // Qp + inner mismatch, origin (IT, JT-32)
        int Q[T + 1][T + 1];    // origin (IT, JT-1); row T = frontier row IT+32
        int QM[T + 1][T + 1];   // origin (IT, JT-1)
        int RQ[T][T + 1];       // Q of tile (J,J)
      } w;
    } u;
    int X[T][T + 1];
    ZuTab E;
    ZuPre pre[2][G][T];
    uint8_t rs[72], cs[72];
    int acc[G][T];
    int16_t list[G * T];
    int nlist[2];  // by band parity
  };

  __device__ static void init(const TriArgs& a, Smem& sm) { load_tab(a, sm.E); }

  // Sequence-only terms of cell (i,j) for the RNAfold tables (ZuPre reused):
  // valid = pair (i, j-1) canonical; r1 = hl_energy(i, j-1); r5c = the stack
  // energy sr(i, j-1); sr = 0 (loops of this grammar add no stacked pair);
  // outI = generic outer mismatch of (i, j-1); boff / in0 = interior
  // mismatches of (i,j) and (i+1,j-1) as inner pairs; f11..f22 the int-table
  // loops minus their inner mismatch (sizes (1,1), (1,2), (2,1), (2,2)).
  __device__ static bool pre(const ZuView& z, int i, int j, ZuPre* o) {
    const int d = j - i, n = z.n;
    o->valid = 0;
    if (j > n || d < 2 || bpc(z.R(i), z.C(j - 1)) == 7) return false;
    o->valid = 1;
    const DevRna* P = z.P;
    const ZuTab& E = z.E;
    const int Ri = z.R(i), Ri1 = z.R(i + 1), Ri2 = z.R(i + 2), Ri3 = z.R(i + 3);
    const int Cj1 = z.C(j - 1), Cj2 = z.C(j - 2), Cj3 = z.C(j - 3), Cj4 = z.C(j - 4);
    const int ocb = bpc(Ri, Cj1);
    // hl_energy(i, j-1) (librna_impl.h:181-193): size d-2
    const int size = d - 2;
    int hp = 600;
    if (size >= 3) {
      const int ent = z.hl[size];  // hairpin[size] up to 30, extrapolated beyond
      const int hmm = P->mmH[ocb][Ri1][Cj2];
      hp = size == 3 ? ent + z.tau(Ri, Cj1) : ent + hmm;
      const int cls = size == 3 ? 0 : size == 4 ? 1 : size == 6 ? 2 : -1;
      if (cls >= 0) {
        const int sv = z.spec(i, cls);
        if (sv != SPEC_NONE) hp = sv;
      }
    }
    int e11 = 0, e12 = 0, e21 = 0, e22 = 0;
    if (d >= 7) {  // (sl, sr) = (1,1): inner pair (i+2, j-3)
      e11 = P->int11[ocb][bpc(Cj3, Ri2)][Ri1][Cj2];
      e12 = P->int21[ocb][bpc(Cj4, Ri2)][Ri1][Cj3][Cj2];
      e21 = P->int21[bpc(Cj3, Ri3)][ocb][Cj2][Ri1][Ri2];
      e22 = P->int22[ocb][bpc(Cj4, Ri3)][Ri1][Ri2][Cj3][Cj2];
    }
    o->r1 = d >= 3 ? hp : 0;
    o->r5c = E.stack[ocb][bpc(Cj2, Ri1)];
    o->sr = 0;
    o->uoff = 0;
    o->boff = (i >= 1 && j <= n - 1) ? z.inner(i, j) : kNoB;
    o->outI = E.mmI[ocb][Ri1][Cj2];
    o->in0 = d >= 4 ? z.inner(i + 1, j - 1) : 0;
    if (d >= 7) {
      o->f11 = (int16_t)(e11 - z.inner(i + 2, j - 2));
      o->f12 = (int16_t)(e12 - z.inner(i + 2, j - 3));
      o->f21 = (int16_t)(e21 - z.inner(i + 3, j - 2));
      o->f22 = (int16_t)(e22 - z.inner(i + 3, j - 3));
    }
    return true;
  }

  // Sequence-only terms of band [ld0, ld0+G) into pre[buf] and the list of
  // cells that can close an interior loop (Qp exists, j - i >= 6), in any
  // order: each wave appends its cells with one LDS atomic on nlist[buf],
  // which must be 0 beforehand.  Called by whole waves, t in [0, 192)
  // (cell t of the band's G*T = 96).
  __device__ static void prep_band(const ZuView& z, int IT, int JT, int ld0, Smem& sm, int t, int buf) {
    bool listed = false;
    const int q = t / T, p = t % T, ld = ld0 + q;
    if (t < G * T) {
      const int rmin = ld < 0 ? -ld : 0, ncell = T - (ld < 0 ? -ld : ld);
      if (ld <= T - 1 && p < ncell)
        listed = pre(z, IT + rmin + p, JT + rmin + p + ld, &sm.pre[buf][q][p]) && ld + JT - IT >= 6;
    }
    const unsigned long long m = __ballot(listed);
    if (m) {
      int base = 0;
      if ((threadIdx.x & 63) == 0) base = atomicAdd(&sm.nlist[buf], __popcll(m));
      base = __builtin_amdgcn_readfirstlane(base);
      if (listed) sm.list[base + __popcll(m & ((1ull << (threadIdx.x & 63)) - 1))] = (int16_t)(q * T + p);
    }
  }

  struct MP {
    static constexpr bool MAXP = false, CHAIN = false, THR = true;
    static constexpr int NEUTRAL = INV, NTv = NT_RF, LT = RT_Q, RT = RT_Q;
  };

  __device__ static void tile(const TriArgs& a, const TriInst& in, int I, int J, Smem& sm, int part = 0, int nparts = 1) {
    const int tid = threadIdx.x, IT = I * T, JT = J * T, n = in.n, nb = in.nb;
    const uint8_t* seq = a.codes + in.seq_off;
    const DevRna* P = (const DevRna*)a.rna;
    const ZuView z{sm.E, P, a.hl_ent, sm.rs, sm.cs, IT - 1, JT - 34, n, a.spec + in.seq_off, a.n_codes};
    for (int e = tid; e < T * (T + 1); e += NTHR) (&sm.X[0][0])[e] = INV;
    for (int e = tid; e < G * T; e += NTHR) (&sm.acc[0][0])[e] = INV;
    if (tid == 0) sm.nlist[0] = 0;
    if (tid < 72) {
      const int x = IT - 1 + tid, y = JT - 34 + tid;
      sm.rs[tid] = (x >= 0 && x < n) ? seq[x] : 0;
      sm.cs[tid] = (y >= 0 && y < n) ? seq[y] : 0;
    }
    // QM split points in blocks strictly between I and J: Q (x) Q
    middle_X<MP, false>(a, in, I, J, sm.u.stage, sm.X, part, nparts);
    auto fetch = [&](int tab, int i, int j) -> int {
      if (i < 0 || j < 0 || i > j || j > n) return INV;
      const int I2 = i / T, J2 = j / T;
      if (I2 >= nb || J2 >= nb) return INV;
      return rec_ptr(a, in, NT_RF, I2, J2)[tab * T * T + (i % T) * T + (j % T)];
    };
    // every global read of the frontier before the first LDS write (one
    // memory latency): B quadrants as in the Zuker tile, Q of tile (J,J),
    // and one halo cell of Q and QM per thread below 65 (row IT+32, then
    // column JT-1)
    const int* qb[3] = {J - 1 >= I ? rec_ptr(a, in, NT_RF, I, J - 1) + RT_B * T * T : nullptr,
                        I + 1 <= J - 1 ? rec_ptr(a, in, NT_RF, I + 1, J - 1) + RT_B * T * T : nullptr,
                        I + 1 <= J ? rec_ptr(a, in, NT_RF, I + 1, J) + RT_B * T * T : nullptr};
    int4 qv[3];
#pragma unroll
    for (int quad = 0; quad < 3; ++quad)
      qv[quad] = qb[quad] ? *(const int4*)(qb[quad] + tid * 4) : make_int4(INV, INV, INV, INV);
    const int4 rq = I < J ? *(const int4*)(rec_ptr(a, in, NT_RF, J, J) + RT_Q * T * T + tid * 4)
                          : make_int4(INV, INV, INV, INV);
    const int hr = tid < T + 1 ? T : tid - (T + 1), hc = tid < T + 1 ? tid : 0;
    int hq = INV, hm = INV;
    if (tid < 2 * T + 1) {
      hq = fetch(RT_Q, IT + hr, JT - 1 + hc);
      hm = fetch(RT_QM, IT + hr, JT - 1 + hc);
    }
    for (int e = tid; e < T * T; e += NTHR) sm.u.w.B[e / T][T + e % T] = INV;
    for (int e = tid; e < (T + 1) * (T + 1); e += NTHR) {
      const int rr = e / (T + 1), cc = e % (T + 1);
      if (rr != T && cc != 0) sm.u.w.Q[rr][cc] = sm.u.w.QM[rr][cc] = INV;
    }
#pragma unroll
    for (int quad = 0; quad < 3; ++quad) {
      const int x = tid * 4, rr = x / T + (quad ? T : 0), cc = x % T + (quad == 2 ? T : 0);
      int* dst = &sm.u.w.B[rr][cc];
      dst[0] = qv[quad].x;
      dst[1] = qv[quad].y;
      dst[2] = qv[quad].z;
      dst[3] = qv[quad].w;
    }
    if (I < J) {
      int* dst = &sm.u.w.RQ[(tid * 4) / T][(tid * 4) % T];
      dst[0] = rq.x;
      dst[1] = rq.y;
      dst[2] = rq.z;
      dst[3] = rq.w;
    }
    if (tid < 2 * T + 1) {
      sm.u.w.Q[hr][hc] = hq;
      sm.u.w.QM[hr][hc] = hm;
    }
    const int* gLQ = rec_ptr(a, in, NT_RF, I, I) + RT_Q * T * T;  // Q tile (I,I), read in place
    const int dlo = I == J ? 0 : -(T - 1);
    __syncthreads();
    if (tid >= 64) prep_band(z, IT, JT, dlo, sm, tid - 64, 0);
    __syncthreads();
    int32_t* gqp = rec_ptr(a, in, NT_RF, I, J) + RT_QP * T * T;
    // interior-loop lane of this thread (sizes capped at 29: row 30 and
    // column 30 of the loop energies masked)
    const int ap = tid & 31, aa = ap + 1;
    int fg[29];
#pragma unroll
    for (int b = 2; b <= 30; ++b) {
      const bool special = ap >= 29 || aa == 1 || (aa == 2 && b <= 3) || (aa == 3 && b == 2) || b == 30;
      fg[b - 2] = special ? GZu::kMask : P->F[ap < 29 ? aa : 1][b];
    }
    const int f_a1 = P->F[ap < 29 ? aa : 1][1], f_1a = P->F[1][ap < 29 ? aa : 1];
    for (int ld0 = dlo, buf = 0; ld0 <= T - 1; ld0 += G, buf ^= 1) {
      const int qn = imin(G, T - ld0);
      int cpre[G + 1];
      cpre[0] = 0;
      for (int q = 0; q < G; ++q) cpre[q + 1] = cpre[q] + (q < qn ? T - abs(ld0 + q) : 0);
      const int nv = sm.nlist[buf], wl = ((nv + 1) >> 1) * 64, total = wl + 2 * cpre[G];
      if (tid == 0) sm.nlist[buf ^ 1] = 0;  // counted by the next band's prep, after the barrier
      for (int w = tid; w < total; w += NTHR) {
        if (w < wl) {
          const int v = w >> 5;
          if (v >= nv) continue;
          const int qp = sm.list[v], q = qp / T, p = qp % T, ld = ld0 + q;
          const int r = (ld < 0 ? -ld : 0) + p, c = r + ld, i = IT + r, j = JT + c;
          // the Zuker loop window of the cell (i-1, j+1) with sizes <= 29
          const int best = GZu::window_lane(z, sm, sm.pre[buf][q][p], ap, i - 1, j + 1, r - 1, c + 1, fg, f_a1, f_1a, 29);
          if (ap == 31) sm.acc[q][p] = best;
        } else {
          // QM split points inside blocks I and J (Q(i,k) and Q(k,j) are
          // empty for j-k < 3 and k-i < 3, so all 32 k of a block are taken)
          const int e = (w - wl) >> 1, part = w & 1;
          int q = 0;
          while (e >= cpre[q + 1]) ++q;
          const int ld = ld0 + q, r = (ld < 0 ? -ld : 0) + (e - cpre[q]), c = r + ld;
          int best = INV;
          if (part == 0) {
            if (I == J) {
#pragma unroll 8
              for (int x = 0; x < T; ++x) best = imin(best, sm.u.w.Q[r][x + 1] + sm.u.w.Q[x][c + 1]);
            } else {
              const int4* g4 = (const int4*)(gLQ + r * T);  // the whole row requested at once
              int4 u4[T / 4];
#pragma unroll
              for (int x4 = 0; x4 < T / 4; ++x4) u4[x4] = g4[x4];
#pragma unroll
              for (int x4 = 0; x4 < T / 4; ++x4) {
                const int4 u = u4[x4];
                const int x = x4 * 4;
                best = imin(imin(best, u.x + sm.u.w.Q[x][c + 1]), u.y + sm.u.w.Q[x + 1][c + 1]);
                best = imin(imin(best, u.z + sm.u.w.Q[x + 2][c + 1]), u.w + sm.u.w.Q[x + 3][c + 1]);
              }
            }
          } else if (I < J) {
#pragma unroll 8
            for (int x = 0; x < T; ++x) best = imin(best, sm.u.w.Q[r][x + 1] + sm.u.w.RQ[x][c]);
          }
          best = imin(best, __builtin_amdgcn_update_dpp(best, best, 0xB1, 0xf, 0xf, false));  // the pair's other lane
          if (part == 0 && best < INV_THR) sm.X[r][c] = imin(sm.X[r][c], best);
        }
      }
      __syncthreads();
      // light phase: wave 0 walks the band's diagonals; waves 1-3 prepare the next band
      if (tid >= 64) {
        if (ld0 + G <= T - 1) prep_band(z, IT, JT, ld0 + G, sm, tid - 64, buf ^ 1);
      } else {
        for (int q = 0; q < qn; ++q) {
          const int ld = ld0 + q, rmin = ld < 0 ? -ld : 0, ncell = T - (ld < 0 ? -ld : ld);
          if (tid < ncell) {
            const int p = tid, r = rmin + p, c = r + ld, i = IT + r, j = JT + c, d = j - i;
            // every operand first (finished diagonals and this cell's own
            // terms; all addresses inside the arrays), then branch-free
            // arithmetic: one LDS round trip per diagonal step
            const ZuPre pc = sm.pre[buf][q][p];
            const int accv = sm.acc[q][p];
            const int Bi = sm.u.w.B[r + 1][c + 31];  // B(i+1, j-1): the stack
            const int mlv = sm.u.w.QM[r + 1][c];     // QM(i+1, j-1): the multiloop
            const int Xv = sm.X[r][c], Ql = sm.u.w.Q[r][c], Qd = sm.u.w.Q[r + 1][c + 1];
            const bool live = i <= j && j <= n, v = live && pc.valid;
            int qp = imin(accv, d >= 3 ? (int)pc.r1 : INV);
            qp = d >= 4 && Bi < INV_THR ? imin(qp, Bi - pc.in0 + pc.r5c) : qp;
            qp = mlv < INV_THR ? imin(qp, mlv) : qp;
            qp = v && qp < INV_THR ? qp : INV;
            const int Bv = qp < INV_THR && pc.boff != kNoB ? qp + pc.boff : INV;
            const int qm = live && Xv < INV_THR && d >= 4 ? Xv : INV;
            int qv = imin(imin(qm, qp), imin(Ql, Qd));
            qv = live && d >= 2 && qv < INV_THR ? qv : INV;
            sm.u.w.B[r][c + 32] = Bv;
            sm.u.w.QM[r][c + 1] = qm;
            sm.u.w.Q[r][c + 1] = qv;
            gqp[r * T + c] = qp;
            sm.acc[q][p] = INV;
          }
          wave_lds_sync();
        }
      }
      __syncthreads();
    }
    int32_t* g = rec_ptr(a, in, NT_RF, I, J);
    for (int e = tid; e < T * T; e += NTHR) {
      const int r = e / T, c = e % T;
      g[RT_Q * T * T + e] = sm.u.w.Q[r][c + 1];
      g[RT_QM * T * T + e] = sm.u.w.QM[r][c + 1];
      g[RT_B * T * T + e] = sm.u.w.B[r][T + c];
    }
  }
};

// ------------------------------------------------------------------ launch
int fill_occupancy(int grammar) {
  int blocks = 0;
  hipError_t e;
  if (grammar == G_NUSSINOV) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, tri_fill<GNu>, NTHR, 0);
  else if (grammar == G_MATCHAIN) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, tri_fill<GMc>, NTHR, 0);
  else if (grammar == G_RNAFOLD) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, tri_fill<GRf>, NTHR, 0);
  else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, tri_fill<GZu>, NTHR, 0);
  return e == hipSuccess ? blocks : 0;
}

int launch_fill(int grammar, const TriArgs& a, const unsigned long long* seeds, int nseed, int grid, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(tri_reset, dim3((max(a.total_items, a.total_tiles) + 255) / 256), dim3(256), 0, st, a, seeds, nseed);
  if (a.spec && a.n_codes > 0)
    hipLaunchKernelGGL(tri_spec, dim3((unsigned)((a.n_codes + 255) / 256)), dim3(256), 0, st, a);
  if (grammar == G_NUSSINOV) hipLaunchKernelGGL(tri_fill<GNu>, dim3(grid), dim3(NTHR), 0, st, a);
  else if (grammar == G_MATCHAIN) hipLaunchKernelGGL(tri_fill<GMc>, dim3(grid), dim3(NTHR), 0, st, a);
  else if (grammar == G_RNAFOLD) hipLaunchKernelGGL(tri_fill<GRf>, dim3(grid), dim3(NTHR), 0, st, a);
  else hipLaunchKernelGGL(tri_fill<GZu>, dim3(grid), dim3(NTHR), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace dpx
