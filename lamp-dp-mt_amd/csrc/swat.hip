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
// tag in every word (persistent kernel, bands claimed in order).
// The traceback keeps one 4-bit code per cell (which of g1 / g2 wins M, and
// whether g1 / g2 extend), 8 steps of a lane per 32-bit word, words of a band
// stored lane-major so each store is a coalesced 256 B row.
//
// Boundary rows travel through a ring of nslots rows (more than the bands in
// flight) of 64-bit words (M, g1, and the tag of the band that reads them),
// so the forward pass needs O(nslots * n2) memory.  Without codes (bt == null) the
// pass keeps every cp_every-th boundary row as a checkpoint instead, and the
// traceback recomputes one segment of cp_every bands at a time from its
// checkpoint (the checkpoint strategy of the reference's design notes,
// docs/report/design.tex:226-262).
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "engine.h"

namespace dpx {

typedef struct { int32_t i, j, rule, pos[2]; } SwTrace;



// DPP controls (wave-wide and 16-lane-row shifts)
constexpr int DPP_WAVE_SHL1 = 0x130;  // lane l <- lane l+1 (lane 63 keeps its old value)
constexpr int DPP_WAVE_SHR1 = 0x138;  // lane l <- lane l-1 (lane 0 keeps its old value)
constexpr int DPP_ROW_SHL0 = 0x100;   // + n: lane l <- lane l+n within its row of 16

// v shifted up one lane; lane 0 takes lane 0 of `lane0`
__device__ __forceinline__ int sw_shr1(int lane0, int v) {
  return __builtin_amdgcn_update_dpp(lane0, v, DPP_WAVE_SHR1, 0xf, 0xf, false);
}
// lane 0 of the result = lane N of v (N < 16)
template <int N>
__device__ __forceinline__ int sw_lane0_from(int v) {
  if constexpr (N == 0) return v;
  else return __builtin_amdgcn_mov_dpp(v, DPP_ROW_SHL0 + N, 0xf, 0xf, true);  // other lanes: don't care
}

// One band of 64 rows per wave (lane r: row i = 64 band + r + 1), swept in
// steps: at step t lane r computes column j = t - r, taking M(i-1,j) and
// g1(i-1,j) from lane r-1 through a DPP wave shift (lane 0: the row above,
// fed from the chunk registers) and keeping M(i,j-1), g2(i,j-1) and
// M(i-1,j-1) in registers.
//
// SW_CH = steps per chunk (16, 32 or 64) = publish granularity: the row
// above is read, and the band's last row written, once per chunk; a band
// lags the one above by 64 + SW_CH steps.  CODES: store the traceback codes
// (A.bt); a separate instance without them keeps the step loop free of a
// store branch.  Per cell the traceback code is 4 bits, shifted into a
// 32-bit word per lane every step (8 steps per word, the first step in the
// top nibble): bit 3 = g1 > pair (M rule 2 unless bit 2), bit 2 = g2 > both
// (M rule 3), bit 1 = g1 extends (rule 6, else 5), bit 0 = g2 extends
// (rule 9, else 8).
// One band of alignment A, by the calling wave (band wave-uniform).
template <int SW_CH, bool CODES>
__device__ __forceinline__ void sw_band(const SwArgs& A, const int band) {
  static_assert(SW_CH % 16 == 0 && SW_CH <= 64, "chunk: whole 16-lane rows of feed, one column per lane");
  constexpr int NQ = SW_CH / 16;
  const int lane = threadIdx.x;
  {
    const int i = band * 64 + lane + 1;
    const bool row_ok = i <= A.n1;
    const int ai = row_ok ? A.a[i - 1] : -1;
    const int W = A.n2 + 1;
    const int match = A.match, mismatch = A.mismatch, open = A.open, ext = A.ext;
    const unsigned long long* above = A.hrow + (size_t)(band % A.nslots) * (W + 1);
    unsigned long long* below = A.hrow + (size_t)((band + 1) % A.nslots) * (W + 1);
    const unsigned long long want_tag = (unsigned long long)(band % SW_TAGS + 1);       // written by band - 1
    const unsigned long long put_tag = (unsigned long long)((band + 1) % SW_TAGS + 1);  // read by band + 1
    uint32_t* btb = A.bt + (size_t)band * A.nwords * 64;
    int2* cpb = (A.cp && (band + 1) % A.cp_every == 0) ? A.cp + (size_t)((band + 1) / A.cp_every) * W : nullptr;
    // column -1 acts as -infinity for M and g2 (NEG + score stays far below
    // 0 and never wraps): then column 0 needs no case of its own -- g2 = 0,
    // M = g1 (rule 2) -- and the trace reads rule 7 for g2 there by itself
    constexpr int NEG = -0x20000000;
    constexpr unsigned VMASK = (1u << SW_VAL_BITS) - 1;
    int Mc = NEG, G1c = 0, G2c = NEG, upPrev = NEG, bj = 0;
    int best = row_ok ? -1 : 0x7fffffff, bestt = 0;  // rows past n1 never improve
    const int jstar = row_ok && i == A.n1 ? A.n2 : -0x40000000;  // column of M(n1, n2) on this lane
    int outv = 0;   // M(n1, n2) when this lane's row is n1
    int capM = 0, capG = 0;  // the band's last row, collected by lane shifts during a chunk
    const int tend = (A.n2 + 64 + 7) & ~7;  // steps of the band, rounded up to whole code words
    // The handoff row is written and read with device-coherent (agent
    // scope) accesses; every word carries the tag of the band that reads
    // it, so the reader checks the words themselves (no separate progress
    // counter, and the writer never waits for its stores).  A chunk's words
    // are loaded one chunk ahead and reloaded only if some tag is stale.
    auto load_row = [&](int t0) -> unsigned long long {
      const int col = t0 + lane;
      return col < W && lane < SW_CH
                 ? __hip_atomic_load(above + col, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                 : want_tag << 54;
    };
    auto load_b = [&](int t0) -> int {
      const int bc = t0 - 1 + lane;  // lane l: b[t0 - 1 + l], the symbol lane 0 needs at step t0 + l
      return (bc >= 0 && bc < A.n2) ? A.b[bc] : 0;
    };
    unsigned long long nxt = load_row(0);
    int nxtB = load_b(0);
    // nothing else is in flight before the loop: the loop's wait for its
    // prefetched words then counts only the loop's own memory operations
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    // The previous chunk's stores -- its traceback code words (kept in
    // registers) and the band's last-row cells it collected -- go out at the
    // start of the next chunk, after that chunk's prefetch, and every chunk
    // issues the same memory instructions (a chunk with nothing to store
    // writes to the band's spare code words and the row's spare word): the
    // compiler's wait for the prefetched words at the top of the next chunk
    // can then leave these stores in flight.
    uint32_t wbuf[SW_CH / 8];
    int pendM = 0, pendG = 0, pend_t0 = -SW_CH, pend_tn = SW_CH;
    const int spare_word = A.nwords - 8;
    auto flush_prev = [&]() {
      if (CODES) {
#pragma unroll
        for (int g = 0; g < SW_CH / 8; ++g) {
          const int w = pend_t0 >= 0 && g * 8 < pend_tn ? ((pend_t0 + g * 8) >> 3) : spare_word + g;
          btb[(size_t)w * 64 + lane] = wbuf[g];
        }
      }
      const int jl = pend_t0 + lane + pend_tn - 127;  // lane l holds step l - 64 + tn of that chunk
      const bool keep = pend_t0 >= 0 && lane >= 64 - pend_tn && jl >= 0 && jl <= A.n2;
      const unsigned long long w = (unsigned long long)((unsigned)pendM & VMASK) |
                                   ((unsigned long long)((unsigned)pendG & VMASK) << SW_VAL_BITS) | (put_tag << 54);
      __hip_atomic_store(below + (keep ? jl : W), w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (cpb && keep) cpb[jl] = make_int2(pendM, pendG);  // read by later launches only
    };
    for (int t0 = 0; t0 < tend; t0 += SW_CH) {
      unsigned long long cur = nxt;
      const int chB = nxtB;
      // the chunk's words must carry this band's tag (lanes past the chunk
      // or the row were filled with it); reload until they do, bounded
      if (__ballot((cur >> 54) != want_tag) != 0) {
        int spin = 0;
        do {
          if (++spin > (1 << 22)) {  // bounded wait: flag the error, stop waiting
            if (lane == 0) A.ctrl[1] = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          cur = load_row(t0);
        } while (__ballot((cur >> 54) != want_tag) != 0);
      }
      // the next chunk's words, in flight while this chunk computes; issued
      // before this chunk's stores, so that waiting for them (loads and
      // stores retire in order) leaves the stores in flight
      if (t0 + SW_CH < tend) {
        nxt = load_row(t0 + SW_CH);
        nxtB = load_b(t0 + SW_CH);
      }
      const int chM = (int)((unsigned)cur & VMASK), chG = (int)((unsigned)(cur >> SW_VAL_BITS) & VMASK);
      // feed registers: lanes 0..15 of fM[q] hold chunk columns 16q..16q+15
      int fM[NQ], fG[NQ], fB[NQ];
      fM[0] = chM;
      fG[0] = chG;
      fB[0] = chB;
#pragma unroll
      for (int q = 1; q < NQ; ++q) {
        const int src = ((lane & 15) + 16 * q) << 2;
        fM[q] = __builtin_amdgcn_ds_bpermute(src, chM);
        fG[q] = __builtin_amdgcn_ds_bpermute(src, chG);
        fB[q] = __builtin_amdgcn_ds_bpermute(src, chB);
      }
      flush_prev();
      const int tn = min(SW_CH, tend - t0);  // a multiple of 8
      // every lane's column of every step of the chunk is inside [0, n2):
      // no masks for the columns left of the matrix, past its end, or M(n1,n2)
      const bool interior = t0 >= 64 && t0 + SW_CH <= A.n2;
      auto run = [&](auto edge) {
        constexpr bool EDGE = decltype(edge)::value;
#pragma unroll
        for (int g = 0; g < SW_CH / 8; ++g) {
          if (g * 8 >= tn) break;
          uint32_t word = 0;
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int s = g * 8 + u;  // step within the chunk
            const int t = t0 + s;
            const int q = s >> 4;
            int lM, lG, lB;
            switch (s & 15) {
#define SW_FEED(N) case N: lM = sw_lane0_from<N>(fM[q]); lG = sw_lane0_from<N>(fG[q]); lB = sw_lane0_from<N>(fB[q]); break;
              SW_FEED(0) SW_FEED(1) SW_FEED(2) SW_FEED(3) SW_FEED(4) SW_FEED(5) SW_FEED(6) SW_FEED(7)
              SW_FEED(8) SW_FEED(9) SW_FEED(10) SW_FEED(11) SW_FEED(12) SW_FEED(13) SW_FEED(14)
              default: lM = sw_lane0_from<15>(fM[q]); lG = sw_lane0_from<15>(fG[q]); lB = sw_lane0_from<15>(fB[q]); break;
#undef SW_FEED
            }
            const int upM = sw_shr1(lM, Mc);
            const int upG = sw_shr1(lG, G1c);
            bj = sw_shr1(lB, bj);
            // g2 = max(0, M(i,j-1)+open, g2(i,j-1)+ext); extension wins only when strictly better
            const int o2 = max(Mc + open, 0), e2 = G2c + ext;
            int g2 = max(o2, e2);
            // g1, from the row above
            const int o1 = max(upM + open, 0), e1 = upG + ext;
            const int g1 = max(o1, e1);
            // M: pair first, then g1, then g2 when strictly greater
            const int dg = upPrev + (ai == bj ? match : mismatch);
            const int m1 = max(dg, g1);
            int M = max(m1, g2);
            // the code bits: x > y is the sign of y - x (every value here
            // lies within +-2^30, so the difference does not wrap), shifted
            // into the word with v_alignbit -- no lane masks kept alive
            word = __builtin_amdgcn_alignbit(word, (unsigned)(dg - g1), 31);  // g1 > pair
            word = __builtin_amdgcn_alignbit(word, (unsigned)(m1 - g2), 31);  // g2 > max(pair, g1)
            word = __builtin_amdgcn_alignbit(word, (unsigned)(o1 - e1), 31);  // g1 extends
            word = __builtin_amdgcn_alignbit(word, (unsigned)(o2 - e2), 31);  // g2 extends
            upPrev = upM;
            const int j = t - lane;
            if (EDGE) {
              // lanes left of the matrix (j < 0) hold column -1: -infinity
              // for M and g2; lanes past n2 and rows past n1 are never read
              if (j < 0) { M = NEG; g2 = NEG; }
              if ((unsigned)j <= (unsigned)A.n2 && M > best) { best = M; bestt = t; }
              if (j == jstar) outv = M;
            } else {
              if (M > best) { best = M; bestt = t; }
            }
            Mc = M;
            G1c = g1;
            G2c = g2;
            // collect lane 63's cells: after the chunk's k-th step its value sits in lane 64 - tn + k
            capM = __builtin_amdgcn_update_dpp(M, capM, DPP_WAVE_SHL1, 0xf, 0xf, false);
            capG = __builtin_amdgcn_update_dpp(g1, capG, DPP_WAVE_SHL1, 0xf, 0xf, false);
          }
          wbuf[g] = word;
        }
      };
      if (interior) run(std::false_type{});
      else run(std::true_type{});
      pendM = capM;
      pendG = capG;
      pend_t0 = t0;
      pend_tn = tn;
    }
    flush_prev();
    if (jstar >= 0) A.out[0] = outv;
    // best cell of this band: largest M, then smallest (i, j)
    if (row_ok && best >= 0) {
      const int bestj = bestt - lane;
      // score in the top 28 bits, then the complement of the 36-bit linear
      // index (the host bounds both): largest score, then smallest (i, j)
      const unsigned long long key = ((unsigned long long)(unsigned)best << SW_KEY_BITS) |
                                     (SW_KEY_MASK - ((unsigned long long)i * (unsigned long long)W + (unsigned long long)bestj));
      atomicMax(A.best, key);
    }
  }
}

// one wave per workgroup: the band claim and every wait are wave-uniform
template <int SW_CH, bool CODES>
__global__ __launch_bounds__(64) void sw_fill(SwArgs A) {
  for (;;) {
    int band = 0;
    if (threadIdx.x == 0) band = atomicAdd(&A.ctrl[0], 1);
    band = __builtin_amdgcn_readfirstlane(band);
    if (band >= A.nbands) break;
    sw_band<SW_CH, CODES>(A, band);
  }
}

// An alignment's arguments read from device memory into scalar registers
// (the calling wave's index is uniform).
__device__ __forceinline__ SwArgs uniform_args(const SwArgs* src) {
  static_assert(sizeof(SwArgs) % 4 == 0, "whole words");
  int w[sizeof(SwArgs) / 4];
  __builtin_memcpy(w, src, sizeof(SwArgs));
#pragma unroll
  for (int k = 0; k < (int)(sizeof(SwArgs) / 4); ++k) w[k] = __builtin_amdgcn_readfirstlane(w[k]);
  SwArgs A;
  __builtin_memcpy(&A, w, sizeof(SwArgs));
  return A;
}

// Several alignments in one persistent launch: the bands of all of them are
// claimed in one order (alignment by alignment, band by band), so a band's
// predecessor was always claimed before it and is running or done.
template <int SW_CH, bool CODES>
__global__ __launch_bounds__(64) void sw_fill_batch(const SwPair* P, const int32_t* band_pair, int total, int* claim) {
  for (;;) {
    int g = 0;
    if (threadIdx.x == 0) g = atomicAdd(claim, 1);
    g = __builtin_amdgcn_readfirstlane(g);
    if (g >= total) break;
    const int p = __builtin_amdgcn_readfirstlane(band_pair[g]);
    const SwArgs A = uniform_args(&P[p].A);
    sw_band<SW_CH, CODES>(A, g - __builtin_amdgcn_readfirstlane(P[p].band0));
  }
}

// Slot 0 = row 0 tagged for band 0; every other slot cleared (tag 0), so no
// word of an earlier launch or an earlier use of the slot passes as fresh.
__device__ __forceinline__ void sw_reset_body(const SwArgs& A, const int2* row0) {
  const int W = A.n2 + 1;
  const size_t total = (size_t)A.nslots * (W + 1);  // each slot: columns 0..n2 and a spare word
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    unsigned long long w = 0;
    if (t < (size_t)W) {
      const int2 v = row0[t];
      w = (unsigned long long)((unsigned)v.x & ((1u << SW_VAL_BITS) - 1)) |
          ((unsigned long long)((unsigned)v.y & ((1u << SW_VAL_BITS) - 1)) << SW_VAL_BITS) | (1ull << 54);
    }
    A.hrow[t] = w;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) { A.ctrl[0] = 0; A.ctrl[1] = 0; *A.best = 0; A.out[0] = 0; }
}
__global__ void sw_reset(SwArgs A, const int2* row0) { sw_reset_body(A, row0); }
// blockIdx.y = alignment; the batch's claim counter is cleared by one thread
__global__ void sw_reset_batch(const SwPair* P, int* claim) {
  const SwPair& q = P[blockIdx.y];
  sw_reset_body(q.A, q.row0);
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *claim = 0;
}

// Traceback: the reference BFS from (i0,j0) (comp1.cu:100-132) is a single
// chain here; one wave walks it, caching a 64-word window of the band's
// traceback words in LDS (the walk only moves to smaller steps within a band).
// Best cell over row 0 (host-computed: value, column) and the device rows
// (largest M, then smallest (i, j)); start of the trace.  out: [0]=axiom
// M(n1,n2), [1]=best, [2]=best_i, [3]=best_j, [4]=trace i0, [5]=trace j0.
__device__ __forceinline__ void sw_pick_body(const SwArgs& A, int row0_best, int row0_bestj, int row0_axiom,
                                             int trace_from, int* out, int* state) {
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
__global__ void sw_pick(SwArgs A, int row0_best, int row0_bestj, int row0_axiom, int trace_from, int* out,
                        int* state) {
  sw_pick_body(A, row0_best, row0_bestj, row0_axiom, trace_from, out, state);
}
__global__ void sw_pick_batch(const SwPair* P) {  // one thread per alignment
  const SwPair& q = P[blockIdx.x];
  sw_pick_body(q.A, q.row0_best, q.row0_bestj, q.row0_axiom, q.trace_from, q.pick, q.state);
}

__device__ __forceinline__ void sw_trace_body(const SwArgs& A, const uint8_t* row0_mrule, const uint8_t* row0_g2rule,
                                              int i_base, SwTrace* tr, int cap, int* state) {
  constexpr int NW = 64;
  __shared__ uint32_t win[NW][64];  // words w0 .. w0+NW-1 of band wb, [word][row]
  const int lane = threadIdx.x;
  int wb = -1, w0 = 0;
  // the walk's state is wave-uniform: scalar registers
  int i = __builtin_amdgcn_readfirstlane(state[0]), j = __builtin_amdgcn_readfirstlane(state[1]);
  int tab = __builtin_amdgcn_readfirstlane(state[2]), n = __builtin_amdgcn_readfirstlane(state[3]);
  bool ok = __builtin_amdgcn_readfirstlane(state[5]) != 0, ended = __builtin_amdgcn_readfirstlane(state[4]) != 0;
  // rule of a cell from its code nibble, 4 bits per nibble value, for the
  // tables M (g1 beats the pair: 2, g2 beats both: 3, else 1), g1 (5 open,
  // 6 extend) and g2 (8 open, 9 extend)
  constexpr unsigned long long RULE_M = [] {
    unsigned long long k = 0;
    for (int x = 0; x < 16; ++x) k |= (unsigned long long)((x & 4) ? 3 : (x & 8) ? 2 : 1) << (4 * x);
    return k;
  }();
  constexpr unsigned long long RULE_G1 = [] {
    unsigned long long k = 0;
    for (int x = 0; x < 16; ++x) k |= (unsigned long long)(5 + ((x >> 1) & 1)) << (4 * x);
    return k;
  }();
  constexpr unsigned long long RULE_G2 = [] {
    unsigned long long k = 0;
    for (int x = 0; x < 16; ++x) k |= (unsigned long long)(8 + (x & 1)) << (4 * x);
    return k;
  }();
  // the move of each rule, 6 bits per rule: bit 0 i decreases, bit 1 j
  // decreases, bits 2-3 the next table (0 M, 1 g1, 2 g2), bit 4 the walk ends
  constexpr unsigned long long MOVES = (0x10ull << 0) | (0x03ull << 6) | (0x04ull << 12) | (0x08ull << 18) |
                                       (0x10ull << 24) | (0x01ull << 30) | (0x05ull << 36) | (0x10ull << 42) |
                                       (0x02ull << 48) | (0x0Aull << 54);
  while (ok && !ended) {
    if (i_base > 0 && i <= i_base) break;  // the rest of the walk is in an earlier segment
    int rule;
    if (i == 0 || (tab > 6 && j == 0)) {  // the first row and column have fixed rules
      if (tab <= 3) rule = j == 0 ? 0 : row0_mrule[j];
      else if (tab <= 6) rule = 4;
      else rule = j == 0 ? 7 : row0_g2rule[j];
      rule = __builtin_amdgcn_readfirstlane(rule);
    } else {
      const int ir = i - i_base - 1;  // row within the codes of A, from 0
      const int b = ir >> 6, r = ir & 63, t = j + r, w = t >> 3;
      if (b != wb || w < w0 || w >= w0 + NW) {
        wb = b;
        w0 = max(0, w - NW + 1);
        // every load issued before the first store (words past the band's
        // last are clamped to it and never selected): one memory latency
        const uint32_t* src = A.bt + (size_t)b * A.nwords * 64 + lane;
        uint32_t v[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) v[k] = src[(size_t)min(w0 + k, A.nwords - 1) * 64];
#pragma unroll
        for (int k = 0; k < NW; ++k) win[k][lane] = v[k];
        __syncthreads();
      }
      // the first step of a word is in its top nibble (see sw_fill)
      const uint32_t word = __builtin_amdgcn_readfirstlane(win[w - w0][r]);
      const uint32_t nib = (word >> (4 * (7 - (t & 7)))) & 15u;
      const unsigned long long rk = tab <= 3 ? RULE_M : tab <= 6 ? RULE_G1 : RULE_G2;
      rule = (int)(rk >> (4 * nib)) & 15;
    }
    if (n >= cap) { ok = false; break; }
    if (lane == 0) tr[n] = SwTrace{i, j, rule, {0, 0}};
    ++n;
    const int mv = (int)(MOVES >> (6 * rule)) & 63;
    i -= mv & 1;
    j -= (mv >> 1) & 1;
    ended = (mv & 16) != 0;
    tab = ended ? tab : (mv & 12) == 0 ? 0 : (mv & 12) == 4 ? 4 : 7;
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
__global__ __launch_bounds__(64) void sw_trace(SwArgs A, const uint8_t* row0_mrule, const uint8_t* row0_g2rule,
                                               int i_base, SwTrace* tr, int cap, int* state) {
  sw_trace_body(A, row0_mrule, row0_g2rule, i_base, tr, cap, state);
}
// one wave per alignment (blockIdx.x); alignments without a trace skip
__global__ __launch_bounds__(64) void sw_trace_batch(const SwPair* P) {
  const SwPair& q = P[blockIdx.x];
  if (q.tr && q.A.bt) sw_trace_body(q.A, q.r0m, q.r0g2, 0, (SwTrace*)q.tr, q.cap, q.state);
}

int sw_launch(const SwArgs& A, const int2* row0, int grid, int chunk, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t words = (size_t)A.nslots * (A.n2 + 2);
  hipLaunchKernelGGL(sw_reset, dim3((unsigned)std::min<size_t>((words + 255) / 256, 4096)), dim3(256), 0, s, A, row0);
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

int sw_batch_launch(const SwPair* P, int npairs, const int32_t* band_pair, int total_bands, int* claim, size_t max_words,
                    bool codes, bool trace, int grid, int chunk, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(sw_reset_batch, dim3((unsigned)std::min<size_t>((max_words + 255) / 256, 1024), npairs), dim3(256),
                     0, s, P, claim);
  auto k = chunk == 16 ? (codes ? sw_fill_batch<16, true> : sw_fill_batch<16, false>)
         : chunk == 64 ? (codes ? sw_fill_batch<64, true> : sw_fill_batch<64, false>)
                       : (codes ? sw_fill_batch<32, true> : sw_fill_batch<32, false>);
  if (total_bands > 0) hipLaunchKernelGGL(k, dim3(grid), dim3(64), 0, s, P, band_pair, total_bands, claim);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sw_batch_finish(const SwPair* P, int npairs, bool trace, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(sw_pick_batch, dim3(npairs), dim3(1), 0, s, P);
  if (trace) hipLaunchKernelGGL(sw_trace_batch, dim3(npairs), dim3(64), 0, s, P);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sw_occupancy() {
  int blocks = 0;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, sw_fill<32, true>, 64, 0) == hipSuccess ? blocks : 0;
}

}  // namespace dpx
