// engine.h -- data layout shared by the host launcher and the GPU kernels.
//
// Triangular grammars (Nussinov, Zuker, matrix chain) fill the cells (i,j),
// 0 <= i <= j <= n, of one or many instances.  The (n+1)x(n+1) upper
// triangle is cut into T x T tiles; tile (I,J), I <= J, holds cells
// i in [I*T, I*T+T), j in [J*T, J*T+T).  Each tile is one record of NT tables
// of T*T int32 (row-major [i%T][j%T]) stored contiguously in HBM, tiles of an
// instance numbered row-major over the upper triangle of tiles.
//
// Scheduling: tile (I,J) with I<J depends on (I,J-1) and (I+1,J) (and
// through them on every tile inside [I..J]); diagonal tiles have no
// dependency.  One persistent kernel pulls tiles from a queue in HBM; a tile
// is appended to the queue when its second predecessor completes (device-scope
// acq/rel counters), so tiles of many instances interleave in order of
// readiness.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dpx {

constexpr int T = 32;           // tile edge
constexpr int NTHR = 256;       // threads per workgroup of the fill kernels
constexpr int NWAVE = NTHR / 64;
constexpr int INV = 1 << 28;    // "no valid candidate" for min-plus tables
constexpr int INV_THR = 1 << 27;
constexpr int NEG = -(1 << 28); // "no candidate" for max-plus tables
constexpr int ZU_EMPTY = 999999;  // the reference's empty value (Zuker.scala:154)
constexpr int SPEC_NONE = -32768; // no special hairpin starts here

enum Grammar { G_NUSSINOV = 0, G_ZUKER = 1, G_MATCHAIN = 2, G_RNAFOLD = 3 };
constexpr int N_GRAMMARS = 4;
// tables per tile record
constexpr int NT_NU = 1;  // s
constexpr int NT_MC = 1;  // M
// cl, U (= dangle-closed cl + ML_intern), ml, ml1, R11 (row prefix-min of U),
// B (= cl + interior mismatch of the pair as the inner pair of a loop)
constexpr int NT_ZU = 6;
enum { ZT_CL = 0, ZT_U = 1, ZT_ML = 2, ZT_ML1 = 3, ZT_R11 = 4, ZT_B = 5 };
// RNAfold grammar: Q, QM, Qp, B (= Qp + interior mismatch of the pair as an inner pair)
constexpr int NT_RF = 4;
enum { RT_Q = 0, RT_QM = 1, RT_QP = 2, RT_B = 3 };

struct TriInst {
  int32_t n;         // sequence length / chain length; cells 0..n
  int32_t nb;        // tiles per side = ceil((n+1)/T)
  int64_t seq_off;   // first base in codes[] (matrix chain: first dim in dims[])
  int64_t tile0;     // record index of tile (0,0)
  int64_t trace_off; // first trace record of this instance
  int32_t trace_cap; // trace records available
  int32_t out;       // result slot
};

struct TriArgs {
  const TriInst* inst;
  int32_t n_inst;
  int32_t total_tiles;
  int32_t total_items;         // queue slots: tiles, plus one more per extra part of a split middle product
  int32_t mp_split;            // 1: the middle products of long tiles are split into part tasks (mp_parts)
  int32_t* mp;                 // split state (mp_split): per tile the part tasks finished [total_tiles], per
                               // tile its first partial slot [total_tiles], then from mp_part_offset the
                               // partial results, T*T int32 per slot
  unsigned long long* queue;   // total_items slots, ~0 = empty
  int32_t* ctrl;               // [0] next slot to claim, [1] next slot to fill
  int32_t* dep;                // completed predecessors per tile
  int32_t* tab;                // tile records
  const uint8_t* codes;        // bases, codes 1..4 (each instance followed by a 0)
  int64_t n_codes;             // bytes of codes[]
  int16_t* spec;               // special-hairpin energies: [3][n_codes] (tri-, tetra-,
                               // hexaloop whose first base is at that position), or SPEC_NONE
  const int32_t* dims;         // matrix chain: rows then cols, n each
  const void* rna;             // DevRna (Zuker)
  const int32_t* hl_ent;       // hairpin entropy by loop size (Zuker)
  int32_t min_loop;            // Nussinov
  // finalize
  int32_t* results;            // per instance
  void* traces;                // dpx_trace_t records
  int32_t* trace_lens;
  int32_t* st;                 // Zuker st column scratch: n+1 per instance at seq_off+q
  int32_t want_trace;
};

__host__ __device__ inline int tile_index(int nb, int I, int J) { return I * nb - (I * (I - 1)) / 2 + (J - I); }
__host__ __device__ inline int64_t tiles_of(int nb) { return (int64_t)nb * (nb + 1) / 2; }
__host__ __device__ inline unsigned long long pack_item(int inst, int I, int J, int part = 0) {
  // I < 4096 (n <= 30000): bits 12..15 of the I field hold the part index
  return ((unsigned long long)(unsigned)inst << 32) | ((unsigned long long)(unsigned)((part << 12) | I) << 16) |
         (unsigned)J;
}
// min over the 32 lanes of each half wave, valid on its lanes 16..31: quad
// permutations and row rotations give every lane its row's minimum, then
// rows 1 and 3 take the last lane of rows 0 and 2 (row_bcast:15).  All DPP:
// no LDS round trip.  The half wave's lanes must all be active.
__device__ __forceinline__ int dpp_half_min(int v) {
  auto mn = [](int x, int y) { return x < y ? x : y; };
  v = mn(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xf, 0xf, false));   // quad_perm [1,0,3,2]
  v = mn(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xf, 0xf, false));   // quad_perm [2,3,0,1]
  v = mn(v, __builtin_amdgcn_update_dpp(v, v, 0x124, 0xf, 0xf, false));  // row_ror:4
  v = mn(v, __builtin_amdgcn_update_dpp(v, v, 0x128, 0xf, 0xf, false));  // row_ror:8
  return mn(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xa, 0xf, false));  // row_bcast:15 into rows 1, 3
}

// int32 index of the first partial result in TriArgs::mp
__host__ __device__ inline int64_t mp_part_offset(int64_t tiles) { return (2 * tiles + 63) & ~(int64_t)63; }
// Part tasks of the middle product of a tile J - I blocks off the diagonal
// when a launch splits them: 1 (not split) below 8 split blocks, else one per
// 4 blocks, at most 8.
__host__ __device__ inline int mp_parts(int d) {
  const int nk = d - 1;
  return nk < 8 ? 1 : (nk / 4 < 8 ? nk / 4 : 8);
}

// Smith-Waterman/Gotoh (swat.hip)
struct SwArgs {
  const uint8_t* a;     // n1 bytes
  const uint8_t* b;     // n2 bytes
  int n1, n2;
  int match, mismatch, open, ext;
  int nbands, nwords;   // nwords = code words per lane per band: sw_code_words(n2)
  int nslots;           // boundary rows in flight: band b reads slot b % nslots, writes (b+1) % nslots
                        // (never a multiple of SW_TAGS, so a slot's last two users have different tags)
  unsigned long long* hrow;  // nslots x (n2+2) words: M | g1 << 27 | tag << 54 of rows 64*band
                             // (columns 0..n2, then one word that absorbs stores of idle lanes)
  int* ctrl;            // [0] band claim counter, [1] band-wait time-out flag
  uint32_t* bt;         // nbands x nwords x 64 traceback codes, or null (none kept)
  int2* cp;             // checkpoint rows: row 64*cp_every*s at cp + s*(n2+1), s >= 1; or null
  int cp_every;         // bands per checkpoint segment
  unsigned long long* best;  // packed (M << 32) | ~(i*(n2+1)+j)
  int* out;             // [0] = M(n1,n2)
};
// Traceback code words per lane of a band: the steps of the band rounded up
// to whole chunks of the largest chunk (64 steps = 8 words), plus 8 words
// that absorb the stores of a chunk with nothing to store.
__host__ __device__ inline int sw_code_words(int n2) { return (((n2 + 64 + 7) / 8 + 7) & ~7) + 8; }
// A boundary-row word: M and g1 (both in [0, 2^27), which the host ensures)
// and the tag (b mod SW_TAGS) + 1 of the band b that reads it; 0 = cleared.
constexpr int SW_VAL_BITS = 27;
constexpr int SW_TAGS = 1023;
// trace state: [0] i, [1] j, [2] table (0 M, 4 g1, 7 g2), [3] records written,
// [4] 1 when the trace has ended, [5] 0 after an overflow
constexpr int SW_STATE = 6;
// best-cell key of an alignment: score << SW_KEY_BITS | (SW_KEY_MASK - i*(n2+1)+j)
constexpr int SW_KEY_BITS = 36;
constexpr unsigned long long SW_KEY_MASK = (1ull << SW_KEY_BITS) - 1;
int sw_launch(const SwArgs& A, const int2* row0, int grid, int chunk, void* stream);  // chunk 16, 32 or 64
int sw_finish_launch(const SwArgs& A, int row0_best, int row0_bestj, int row0_axiom, int trace_from, int* out,
                     const uint8_t* r0m, const uint8_t* r0g2, void* tr, int cap, int* state, void* stream);
// Traceback through the codes of rows i_base+1 .. i_base+A.n1 (A covers those
// rows), from the state to the first cell in row <= i_base (i_base > 0) or to
// its end (i_base = 0, where row 0's rules apply).
int sw_trace_launch(const SwArgs& A, const uint8_t* r0m, const uint8_t* r0g2, int i_base, void* tr, int cap,
                    int* state, void* stream);
// Alignments batched into one launch: per alignment (device memory) its
// arguments, its first band in the launch's claim order and what the reset,
// pick and traceback kernels read.
struct SwPair {
  SwArgs A;
  int band0;
  int row0_best, row0_bestj, row0_axiom, trace_from, cap;
  const int2* row0;                     // row 0 of M and g1
  const uint8_t* r0m;                   // row 0's M rules
  const uint8_t* r0g2;                  // row 0's g2 rules
  int* pick;                            // 6 ints (sw_pick's out)
  int* state;                           // SW_STATE ints
  void* tr;                             // cap trace records, or null
};
// reset + persistent fill over all bands of the batch (codes: every pair
// keeps traceback codes); then best cell, start and traceback per pair
int sw_batch_launch(const SwPair* P, int npairs, const int32_t* band_pair, int total_bands, int* claim, size_t max_words,
                    bool codes, bool trace, int grid, int chunk, void* stream);
int sw_batch_finish(const SwPair* P, int npairs, bool trace, void* stream);
int sw_occupancy();

// launchers (tri.hip, finalize.hip); stream is a hipStream_t
int fill_occupancy(int grammar);
int launch_fill(int grammar, const TriArgs& a, const unsigned long long* seeds, int nseed, int grid, void* stream);
int launch_finalize(int grammar, const TriArgs& a, int count, int n_max, void* stream);

}  // namespace dpx
