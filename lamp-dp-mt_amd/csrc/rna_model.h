// rna_model.h -- Turner energy model as the GPU kernels consume it.
//
// The host reads a ViennaRNA v2.0 .par file (rna_params.cpp) into the same
// scaled values as the reference's get_scaled_parameters at 37 C
// (src/librna/vienna/vienna.c:13-96), then flattens them into DevRna: every
// table the Zuker grammar reads, plus precomputed loop-length terms so that no
// floating point runs on the GPU (the reference's only FP on the path is
// lxc*log(l/30), librna_impl.h:149, evaluated here on the host with libm).
#pragma once
#include <stdint.h>

namespace dpx {

constexpr int kMaxLoop = 30;     // energy_const.h:27 MAXLOOP
constexpr int kNoBp = 7;         // librna_impl.h:92 NO_BP

// Device-side energy tables.  Base codes A=1 C=2 G=3 U=4 (0 unused); pair
// codes CG=1 GC=2 GU=3 UG=4 AU=5 UA=6, 7 = not a canonical pair
// (librna_impl.h:90-92,106-114).
struct DevRna {
  int32_t stack[8][8];
  int32_t mmH[8][5][5], mmI[8][5][5], mm1nI[8][5][5], mm23I[8][5][5], mmM[8][5][5], mmExt[8][5][5];
  int32_t d5[8][5], d3[8][5];
  int32_t hairpin[31], bulge[31];
  int32_t il_ent[61];            // interior-loop entropy for total size 0..60
  int16_t F[31][31];             // il_ent(a+b) + asymmetry(a,b), 1<=a,b<=30
  int32_t ninio0, ninio_max, il5, ml_closing, ml_intern, term_au;
  int32_t n_loop[3];             // entries of triloops, tetraloops, hexaloops
  int32_t loop_key[3][40];       // 3 bits per base (N=0..U=4); -1 never matches
  int32_t loop_e[3][40];
  int16_t int11[8][8][5][5];
  int16_t int21[8][8][5][5][5];
  int16_t int22[8][8][5][5][5][5];
};

struct HostRna {
  DevRna dev;
  double lxc;
  int16_t hairpin30;             // hairpin[MAXLOOP], base of the long-hairpin entropy
  // hairpin entropy for loop size l (any l): the host table the device uses.
  int32_t hl_ent(int l) const;
};

// Returns 0 on success, else an error code; msg receives a description.
int load_par_file(const char* path, HostRna* out, char* msg, int msg_len);

}  // namespace dpx
