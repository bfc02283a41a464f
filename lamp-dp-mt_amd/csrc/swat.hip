// swat.hip -- Smith-Waterman/Gotoh two-track grammar on gfx950.
//
// Recurrences (bench/dp_swat/comp1.cu:50-97, SWatAffine.scala:22-69) with
// parametric scores; per cell the reference evaluates g2, g1, then M and keeps
// the first candidate, later ones only when strictly greater:
//   g2(i,j) = j==0 ? 0 [7] : max{ max(0,M(i,j-1)+open) [8], max(0,g2(i,j-1)+ext) [9] }
//   g1(i,j) = i==0 ? 0 [4] : max{ max(0,M(i-1,j)+open) [5], max(0,g1(i-1,j)+ext) [6] }
//   M(i,j)  = max{ 0 [0, i=j=0]; M(i-1,j-1)+(a==b?match:mismatch) [1]; g1 [2]; g2 [3] }
//
// Layout: rows are cut into bands of 64; one wave64 owns a band, lane r owns
// row i = 64*band + r + 1 and sweeps the columns skewed by r (anti-diagonal
// wavefront in registers): at step t lane r computes column j = t - r, takes
// M(i-1,j), g1(i-1,j) from lane r-1 through a DPP wave shift and keeps its own
// M(i,j-1), g2(i,j-1) and the diagonal M(i-1,j-1) in registers.  The band's
// last row goes to HBM and is published to the next band in chunks through a
// device-scope progress counter (persistent kernel, bands claimed in order).
// The traceback keeps one 4-bit code per cell (M rule, g1 ext?, g2 ext?),
// 8 steps of a lane per 32-bit word, words of a band stored lane-major so each
// store is a coalesced 256 B row.
//
// Boundary rows travel through a ring of nslots rows (more than the bands in
// flight), each with a progress counter tagged by its writer band, so the
// forward pass needs O(nslots * n2) memory.  Without codes (bt == null) the
// pass keeps every cp_every-th boundary row as a checkpoint instead, and the
// traceback recomputes one segment of cp_every bands at a time from its
// checkpoint (the checkpoint strategy of the reference's design notes,
// docs/report/design.tex:226-262).
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "engine.h"

namespace dpx {

typedef struct { int32_t i, j, rule, pos[2]; } SwTrace;



__device__ __forceinline__ int sw_shr1(int lane0_val, int v) {
  // lanes 1..63 receive lane-1's v; lane 0 keeps lane0_val (DPP wave_shr:1)
  return __builtin_amdgcn_update_dpp(lane0_val, v, 0x138, 0xf, 0xf, false);
}

// SW_CH = steps per chunk = publish granularity (columns), a multiple of 8 up
// to 64: a band lags the one above by 64 + SW_CH steps, and waits for it (or
// publishes to the band below) once per chunk.  Variants 16, 32, 64; the
// tuning table picks one per length bucket.  CODES: store the traceback codes
// (A.bt); a separate instance without them keeps the step loop free of a
// store branch, which the compiler would otherwise hoist by duplicating the
// loop at twice the registers.
template <int SW_CH, bool CODES>
__global__ __launch_bounds__(64) void sw_fill(SwArgs A) {
  static_assert(SW_CH % 8 == 0 && SW_CH <= 64, "chunk: whole code words, one column per lane");
  const int lane = threadIdx.x;
  for (;;) {
    // one wave per workgroup: the band claim and every wait are wave-uniform
    int band = 0;
    if (lane == 0) band = atomicAdd(&A.ctrl[0], 1);
    band = __builtin_amdgcn_readfirstlane(band);
    if (band >= A.nbands) break;
    const int i = band * 64 + lane + 1;
    const bool row_ok = i <= A.n1;
    const int ai = row_ok ? A.a[i - 1] : -1;
    const int W = A.n2 + 1;
    const int2* above = A.hrow + (size_t)(band % A.nslots) * W;
    int2* below = A.hrow + (size_t)((band + 1) % A.nslots) * W;
    const unsigned long long* pa = A.prog + band % A.nslots;
    unsigned long long* pb = A.prog + (band + 1) % A.nslots;
    const unsigned long long tag_above = (unsigned long long)band << 32;  // written by band - 1
    uint32_t* btb = A.bt + (size_t)band * A.nwords * 64;
    int2* cpb = (A.cp && (band + 1) % A.cp_every == 0) ? A.cp + (size_t)((band + 1) / A.cp_every) * W : nullptr;
    // column -1 acts as -infinity for M and g2 (NEG + score stays far below
    // 0 and never wraps): then column 0 needs no case of its own -- g2 = 0,
    // M = g1 (rule 2) -- and the trace reads rule 7 for g2 there by itself
    constexpr int NEG = -0x20000000;
    int Mc = NEG, G1c = 0, G2c = NEG, upPrev = NEG;
    int best = row_ok ? -1 : 0x7fffffff, bestj = 0;  // rows past n1 never improve
    const int jstar = row_ok && i == A.n1 ? A.n2 : -0x40000000;  // column of M(n1, n2) on this lane
    int known = 0;  // columns of `above` known to be published
    int chM = 0, chG = 0, chB = 0;  // chunk of the row above and of b, one column per lane
    int bj = 0;  // b[j-1] of this lane's column j = t - lane, shifted in from lane 0
    int outv = 0;  // M(n1, n2) when this lane's row is n1
    const int tend = (A.n2 + 64 + 7) & ~7;  // steps of the band, rounded up to whole code words
    for (int t0 = 0; t0 < tend; t0 += SW_CH) {
      // make the next chunk of the row above available, then load it; the
      // steps of the chunk below issue no loads, so no step waits on memory
      const int need = min(t0 + SW_CH, W);
      if (known < need) {
        // columns published by band - 1 (an older tag is an earlier use of the slot)
        auto cols = [&]() {
          const unsigned long long v = __hip_atomic_load(pa, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const int c = (v >> 32) == (tag_above >> 32) ? (int)(uint32_t)v : 0;
          return __builtin_amdgcn_readfirstlane(c);
        };
        int k = cols();
        for (int spin = 0; k < need; ++spin) {
          if (spin > (1 << 24)) {  // bounded wait: flag the error, stop the band
            if (lane == 0) A.ctrl[1] = 1;
            k = need;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          k = cols();
        }
        known = k;
      }
      // The handoff row is written and read with device-coherent (agent
      // scope) accesses, so neither side needs a cache writeback/invalidate:
      // the producer waits for its row stores before it raises the counter.
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      const int col = t0 + lane;
      if (col < W) {
        const unsigned long long v = __hip_atomic_load((const unsigned long long*)(above + col), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
        chM = (int)(uint32_t)v;
        chG = (int)(uint32_t)(v >> 32);
      }
      // lane l holds b[t0 - 1 + l]: the symbol lane 0 needs at step t0 + l
      const int bc = t0 - 1 + lane;
      chB = (bc >= 0 && bc < A.n2) ? A.b[bc] : 0;
      const int tn = min(SW_CH, tend - t0);  // a multiple of 8
      int sM = 0, sG = 0;  // lane l keeps (M, g1) of the band's last row at step t0 + l
      for (int tc = 0; tc < tn; tc += 8) {
        uint32_t word = 0;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int t = t0 + tc + u;
          // chunk values of this step, broadcast from lane tc+u (wave-uniform index)
          const int hM = __builtin_amdgcn_readlane(chM, tc + u), hG = __builtin_amdgcn_readlane(chG, tc + u);
          bj = sw_shr1(__builtin_amdgcn_readlane(chB, tc + u), bj);
          const int upM = sw_shr1(hM, Mc);
          const int upG = sw_shr1(hG, G1c);
          const int j = t - lane;
          // g2 (j == 0: both candidates are 0, rule 7 is the trace's)
          const int o = max(0, Mc + A.open), e = max(0, G2c + A.ext);
          const int r2 = e > o;
          const int g2 = r2 ? e : o;
          // g1
          const int o1 = max(0, upM + A.open), e1 = max(0, upG + A.ext);
          const int r1 = e1 > o1;
          const int g1 = r1 ? e1 : o1;
          // M (j == 0: the diagonal comes from column -1, so g1 wins, rule 2)
          const int dg = upPrev + (ai == bj ? A.match : A.mismatch);
          int M = dg, rm = 1;
          if (g1 > M) { M = g1; rm = 2; }
          if (g2 > M) { M = g2; rm = 3; }
          upPrev = upM;
          // Lanes left of the matrix (j < 0) hold column -1: -infinity for M
          // and g2; lanes past n2 and rows past n1 are never read by a valid cell.
          Mc = j < 0 ? NEG : M;
          G1c = g1;
          G2c = j < 0 ? NEG : g2;
          if ((unsigned)j <= (unsigned)A.n2 && M > best) { best = M; bestj = j; }
          if (j == jstar) outv = M;
          const int lastM = __builtin_amdgcn_readlane(M, 63), lastG = __builtin_amdgcn_readlane(g1, 63);
          if (lane == tc + u) { sM = lastM; sG = lastG; }
          word |= (uint32_t)(rm | (r1 << 2) | (r2 << 3)) << (4 * u);
        }
        if (CODES) btb[(size_t)((t0 + tc) >> 3) * 64 + lane] = word;
      }
      // the chunk's cells of the band's last row, one column per lane
      {
        const int jl = t0 + lane - 63;
        if (lane < tn && jl >= 0 && jl <= A.n2) {
          __hip_atomic_store((unsigned long long*)(below + jl),
                             (unsigned long long)(uint32_t)sM | ((unsigned long long)(uint32_t)sG << 32),
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (cpb) cpb[jl] = make_int2(sM, sG);  // read by later launches only
        }
      }
      // publish the finished columns of this band's last row
      const int t = t0 + tn - 1;
      const int done = min(t - 63 + 1, W);  // columns 0..t-63 of row 64*band+64
      if (done > 0) {
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this band's row stores are complete
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (lane == 63)
          __hip_atomic_store(pb, ((unsigned long long)(band + 1) << 32) | (unsigned)done, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if (jstar >= 0) A.out[0] = outv;
    // best cell of this band: largest M, then smallest (i, j)
    if (row_ok && best >= 0) {
      // score in the top 28 bits, then the complement of the 36-bit linear
      // index (the host bounds both): largest score, then smallest (i, j)
      const unsigned long long key = ((unsigned long long)(unsigned)best << SW_KEY_BITS) |
                                     (SW_KEY_MASK - ((unsigned long long)i * (unsigned long long)W + (unsigned long long)bestj));
      atomicMax(A.best, key);
    }
  }
}

__global__ void sw_reset(SwArgs A, const int2* row0) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int W = A.n2 + 1;
  if (t < W) A.hrow[t] = row0[t];  // slot 0: row 0 (tag 0: "written by band -1")
  if (t < A.nslots) A.prog[t] = t == 0 ? (unsigned long long)W : 0ull;
  if (t == 0) { A.ctrl[0] = 0; A.ctrl[1] = 0; *A.best = 0; A.out[0] = 0; }
}

// Traceback: the reference BFS from (i0,j0) (comp1.cu:100-132) is a single
// chain here; one wave walks it, caching a 64-word window of the band's
// traceback words in LDS (the walk only moves to smaller steps within a band).
// Best cell over row 0 (host-computed: value, column) and the device rows
// (largest M, then smallest (i, j)); start of the trace.  out: [0]=axiom
// M(n1,n2), [1]=best, [2]=best_i, [3]=best_j, [4]=trace i0, [5]=trace j0.
__global__ void sw_pick(SwArgs A, int row0_best, int row0_bestj, int row0_axiom, int trace_from, int* out,
                        int* state) {
  const unsigned long long key = *A.best;
  const int W = A.n2 + 1;
  int bv = row0_best, bi = 0, bj = row0_bestj;
  if (A.nbands > 0) {
    const int dv = (int)(key >> SW_KEY_BITS);
    const unsigned long long lin = SW_KEY_MASK - (key & SW_KEY_MASK);
    if (dv > bv) { bv = dv; bi = (int)(lin / (uint32_t)W); bj = (int)(lin % (uint32_t)W); }
  }
  out[0] = A.n1 == 0 ? row0_axiom : A.out[0];
  out[1] = bv; out[2] = bi; out[3] = bj;
  out[4] = trace_from ? bi : A.n1;
  out[5] = trace_from ? bj : A.n2;
  state[0] = out[4];
  state[1] = out[5];
  state[2] = 0;
  state[3] = 0;
  state[4] = 0;
  state[5] = 1;
}

__global__ __launch_bounds__(64) void sw_trace(SwArgs A, const uint8_t* row0_mrule, const uint8_t* row0_g2rule,
                                               int i_base, SwTrace* tr, int cap, int* state) {
  constexpr int NW = 64;
  __shared__ uint32_t win[NW][64];
  const int lane = threadIdx.x;
  int wb = -1, w0 = 0;
  int i = state[0], j = state[1], tab = state[2], n = state[3];
  bool ok = state[5] != 0, ended = state[4] != 0;
  while (ok && !ended) {
    if (i_base > 0 && i <= i_base) break;  // the rest of the walk is in an earlier segment
    int rule;
    if (tab <= 3) {
      if (i == 0 && j == 0) rule = 0;
      else if (i == 0) rule = row0_mrule[j];
      else rule = -1;
    } else if (tab <= 6) {
      rule = i == 0 ? 4 : -1;
    } else {
      rule = j == 0 ? 7 : (i == 0 ? row0_g2rule[j] : -1);
    }
    if (rule < 0) {
      const int ir = i - i_base;  // row within the codes of A
      const int b = (ir - 1) / 64, r = (ir - 1) % 64, t = j + r, w = t >> 3;
      if (b != wb || w < w0 || w >= w0 + NW) {
        wb = b;
        w0 = max(0, w - NW + 1);
        for (int k = 0; k < NW; ++k)
          if (w0 + k < A.nwords) win[k][lane] = A.bt[((size_t)b * A.nwords + w0 + k) * 64 + lane];
        __syncthreads();
      }
      const uint32_t nib = (win[w - w0][r] >> (4 * (t & 7))) & 15u;
      rule = tab <= 3 ? (int)(nib & 3u) : tab <= 6 ? 5 + (int)((nib >> 2) & 1u) : 8 + (int)((nib >> 3) & 1u);
    }
    if (n >= cap) { ok = false; break; }
    if (lane == 0) tr[n] = SwTrace{i, j, rule, {0, 0}};
    ++n;
    if (rule == 1) { --i; --j; tab = 0; }
    else if (rule == 2) { tab = 4; }
    else if (rule == 3) { tab = 7; }
    else if (rule == 5) { --i; tab = 0; }
    else if (rule == 6) { --i; tab = 4; }
    else if (rule == 8) { --j; tab = 0; }
    else if (rule == 9) { --j; tab = 7; }
    else ended = true;
  }
  if (lane == 0) {
    state[0] = i;
    state[1] = j;
    state[2] = tab;
    state[3] = n;
    state[4] = ended ? 1 : 0;
    state[5] = ok ? 1 : 0;
  }
}

int sw_launch(const SwArgs& A, const int2* row0, int grid, int chunk, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const int W = A.n2 + 1;
  const int nthr = max(W, A.nslots);
  hipLaunchKernelGGL(sw_reset, dim3((nthr + 255) / 256), dim3(256), 0, s, A, row0);
  const bool codes = A.bt != nullptr;
  auto k = chunk == 16 ? (codes ? sw_fill<16, true> : sw_fill<16, false>)
         : chunk == 64 ? (codes ? sw_fill<64, true> : sw_fill<64, false>)
                       : (codes ? sw_fill<32, true> : sw_fill<32, false>);
  hipLaunchKernelGGL(k, dim3(grid), dim3(64), 0, s, A);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sw_finish_launch(const SwArgs& A, int row0_best, int row0_bestj, int row0_axiom, int trace_from, int* out,
                     const uint8_t* r0m, const uint8_t* r0g2, void* tr, int cap, int* state, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(sw_pick, dim3(1), dim3(1), 0, s, A, row0_best, row0_bestj, row0_axiom, trace_from, out, state);
  if (tr && A.bt) hipLaunchKernelGGL(sw_trace, dim3(1), dim3(64), 0, s, A, r0m, r0g2, 0, (SwTrace*)tr, cap, state);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sw_trace_launch(const SwArgs& A, const uint8_t* r0m, const uint8_t* r0g2, int i_base, void* tr, int cap,
                    int* state, void* stream) {
  hipLaunchKernelGGL(sw_trace, dim3(1), dim3(64), 0, (hipStream_t)stream, A, r0m, r0g2, i_base, (SwTrace*)tr, cap,
                     state);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sw_occupancy() {
  int blocks = 0;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, sw_fill<32, true>, 64, 0) == hipSuccess ? blocks : 0;
}

}  // namespace dpx
