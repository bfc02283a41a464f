// rna_params.cpp -- ViennaRNA v2.0 parameter-file reader for the GPU model.
//
// Reproduces the values the reference obtains from read_parameter_file +
// get_scaled_parameters (src/librna/vienna/vienna.c:13-96, 171-310) at 37 C:
//  * a section "# name" is followed by value lines; values are read in runs,
//    each run starting on a fresh line and dropping what is left of its last
//    line (get_array1, vienna.c:232-272); a run covers the innermost
//    dimension when the section has index shifts, else the whole block;
//  * tokens: '*' keeps the previous value, 'x' extrapolates
//    v[last] + (int)(0.5 + lxc37*log(i/last)), DEF=-50, INF=INT_MAX/10, NST=0;
//    one /* comment */ per line is removed;
//  * special hairpin lists (Triloops/Tetraloops/Hexaloops) are read until a
//    line without "seq e37 edH"; the line that ends the list is consumed;
//  * at 37 C the rescaling is the identity on the 37 C values, which the
//    reference then stores as short (truncating through int), clamping
//    mismatch_multi, mismatch_exterior and dangles to <= 0.
#include "rna_model.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

namespace dpx {

namespace {

constexpr int kInf = 0x7fffffff / 10;   // energy_const.h:23 INF
constexpr double kLxc37 = 107.856;      // energy_data.h:53 (absent from .par files)

struct ParReader {
  FILE* fp = nullptr;
  bool failed = false;
  std::string err;

  bool line(std::string* out) {
    out->clear();
    char buf[4096];
    bool any = false;
    while (fgets(buf, sizeof buf, fp)) {
      any = true;
      size_t l = strlen(buf);
      if (l && buf[l - 1] == '\n') { buf[l - 1] = 0; out->append(buf); return true; }
      out->append(buf);
    }
    return any;
  }

  // One run of `count` values into dst (vienna.c:232-272 semantics).
  void run(int* dst, int count) {
    int filled = 0, last = 0;
    std::string ln;
    while (filled < count && !failed) {
      if (!line(&ln)) { failed = true; err = "unexpected end of parameter file"; return; }
      size_t c1 = ln.find("/*");
      if (c1 != std::string::npos) {
        size_t c2 = ln.find("*/", c1);
        if (c2 == std::string::npos) { failed = true; err = "unclosed comment"; return; }
        ln.erase(c1, c2 + 2 - c1);
      }
      const char* p = ln.c_str();
      char tok[16];
      int used = 0;
      while (filled < count && sscanf(p, "%15s%n", tok, &used) == 1) {
        p += used;
        int v;
        if (tok[0] == '*') { ++filled; continue; }
        if (tok[0] == 'x') {
          if (filled == 0) { failed = true; err = "cannot extrapolate the first value"; return; }
          v = dst[last] + (int)(0.5 + kLxc37 * log((double)filled / (double)last));
        } else if (!strcmp(tok, "DEF")) {
          v = -50;
        } else if (!strcmp(tok, "INF")) {
          v = kInf;
        } else if (!strcmp(tok, "NST")) {
          v = 0;
        } else {
          char* end = nullptr;
          long lv = strtol(tok, &end, 10);
          if (end == tok) { failed = true; err = std::string("bad token '") + tok + "'"; return; }
          v = (int)lv;
          last = filled;
        }
        dst[filled++] = v;
      }
    }
  }

  // A section of shape dims[0..nd) whose leading indices start at lo[] and
  // stop before dims[]-hi[]; runs are the innermost dimension unless no
  // index is shifted, in which case the whole block is one run.
  void block(int* dst, int nd, const int* dims, const int* lo, const int* hi) {
    bool shifted = false;
    for (int d = 0; d < nd; ++d) shifted |= (lo[d] | hi[d]) != 0;
    long stride = 1;
    for (int d = 1; d < nd; ++d) stride *= dims[d];
    if (!shifted) { run(dst, (int)(stride * dims[0])); return; }
    if (nd == 1) { run(dst + lo[0], dims[0] - lo[0] - hi[0]); return; }
    for (int q = lo[0]; q < dims[0] - hi[0] && !failed; ++q) block(dst + q * stride, nd - 1, dims + 1, lo + 1, hi + 1);
  }

  // Special hairpin list: str gets each entry at stride*q followed by a blank.
  void loops(char* str, int cap, int* e37, int width, int stride) {
    memset(str, 0, cap);
    for (int q = 0; q < 40; ++q) e37[q] = 0;
    std::string ln;
    char fmt[24];
    snprintf(fmt, sizeof fmt, "%%%ds %%d %%d", width);
    for (int q = 0; q < 40; ++q) {
      if (!line(&ln)) break;
      int dh = 0;
      int got = sscanf(ln.c_str(), fmt, str + stride * q, &e37[q], &dh);
      size_t l = strlen(str);
      if (l + 1 < (size_t)cap) { str[l] = ' '; str[l + 1] = 0; }
      if (got != 3) break;
    }
  }
};

inline int16_t as_short(int v) { return (int16_t)v; }  // x86-64 double->short goes through int
inline int clamp0(int v) { return v > 0 ? 0 : v; }

int loop_key(const char* p, int len) {
  static const char kChr[] = "NACGU";
  int key = 0;
  for (int m = 0; m < len; ++m) {
    const char* f = p[m] ? strchr(kChr, p[m]) : nullptr;
    if (!f) return -1;
    key = key * 8 + (int)(f - kChr);
  }
  return key;
}

}  // namespace

int32_t HostRna::hl_ent(int l) const {
  if (l <= kMaxLoop) return dev.hairpin[l];
  return hairpin30 + (int)(lxc * log((l) * (1.0 / kMaxLoop)));
}

int load_par_file(const char* path, HostRna* out, char* msg, int msg_len) {
  auto fail = [&](int code, const std::string& m) {
    if (msg && msg_len > 0) snprintf(msg, msg_len, "%s", m.c_str());
    return code;
  };
  ParReader rd;
  rd.fp = fopen(path, "r");
  if (!rd.fp) return fail(1, std::string("cannot open ") + path);
  std::string ln;
  if (!rd.line(&ln) || ln.compare(0, 30, "## RNAfold parameter file v2.0") != 0) {
    fclose(rd.fp);
    return fail(2, "missing v2.0 header");
  }

  // raw 37 C arrays, shaped as in the reference's energy_data.h
  static int stack[8][8], mm[6][8][5][5], dng[2][8][5], i11[8][8][5][5], i21[8][8][5][5][5], i22[8][8][5][5][5][5];
  static int hp[31], bu[31], in[31], mlp[6], nin[3], misc[4], junk[8 * 8 * 625];
  static char tri[241], tetra[281], hexa[361];
  static int tri_e[40], tetra_e[40], hexa_e[40];
  memset(stack, 0, sizeof stack); memset(mm, 0, sizeof mm); memset(dng, 0, sizeof dng);
  memset(i11, 0, sizeof i11); memset(i21, 0, sizeof i21); memset(i22, 0, sizeof i22);

  const int d_stack[2] = {8, 8}, lo11[2] = {1, 1}, z2[2] = {0, 0};
  const int d_mm[3] = {8, 5, 5}, lo_mm[3] = {1, 0, 0}, z3[3] = {0, 0, 0};
  const int d_dg[2] = {8, 5}, lo_dg[2] = {1, 0};
  const int d_11[4] = {8, 8, 5, 5}, lo_11[4] = {1, 1, 0, 0}, z4[4] = {0, 0, 0, 0};
  const int d_21[5] = {8, 8, 5, 5, 5}, lo_21[5] = {1, 1, 0, 0, 0}, z5[5] = {0, 0, 0, 0, 0};
  const int d_22[6] = {8, 8, 5, 5, 5, 5}, lo_22[6] = {1, 1, 1, 1, 1, 1}, hi_22[6] = {1, 1, 0, 0, 0, 0};
  static const char* kMM[6] = {"mismatch_hairpin", "mismatch_interior", "mismatch_interior_1n",
                               "mismatch_interior_23", "mismatch_multi", "mismatch_exterior"};
  unsigned seen = 0;  // bit per required section
  while (!rd.failed && rd.line(&ln)) {
    char id[256];
    if (sscanf(ln.c_str(), "# %255s", id) != 1) continue;
    std::string s(id);
    if (s == "END") break;
    bool enth = s.size() > 11 && s.compare(s.size() - 11, 11, "_enthalpies") == 0;
    std::string base = enth ? s.substr(0, s.size() - 11) : s;
    int* tgt = nullptr;
    int bit = -1;
    if (base == "stack") { tgt = enth ? junk : &stack[0][0]; rd.block(tgt, 2, d_stack, lo11, z2); bit = 0; }
    else if (base.compare(0, 8, "mismatch") == 0) {
      int k = -1;
      for (int q = 0; q < 6; ++q) if (base == kMM[q]) k = q;
      if (k < 0) { rd.failed = true; rd.err = "unknown section " + s; break; }
      tgt = enth ? junk : &mm[k][0][0][0];
      rd.block(tgt, 3, d_mm, lo_mm, z3);
      bit = 1 + k;
    } else if (base == "dangle5" || base == "dangle3") {
      int k = base == "dangle5" ? 0 : 1;
      tgt = enth ? junk : &dng[k][0][0];
      rd.block(tgt, 2, d_dg, lo_dg, z2);
      bit = 7 + k;
    } else if (base == "int11") { tgt = enth ? junk : &i11[0][0][0][0]; rd.block(tgt, 4, d_11, lo_11, z4); bit = 9; }
    else if (base == "int21") { tgt = enth ? junk : &i21[0][0][0][0][0]; rd.block(tgt, 5, d_21, lo_21, z5); bit = 10; }
    else if (base == "int22") { tgt = enth ? junk : &i22[0][0][0][0][0][0]; rd.block(tgt, 6, d_22, lo_22, hi_22); bit = 11; }
    else if (base == "hairpin") { rd.run(enth ? junk : hp, 31); bit = 12; }
    else if (base == "bulge") { rd.run(enth ? junk : bu, 31); bit = 13; }
    else if (base == "interior") { rd.run(enth ? junk : in, 31); bit = 14; }
    else if (s == "ML_params") { rd.run(mlp, 6); bit = 15; }
    else if (s == "NINIO") { rd.run(nin, 3); bit = 16; }
    else if (s == "Misc") { rd.run(misc, 4); bit = 17; }
    else if (s == "Triloops") { rd.loops(tri, 241, tri_e, 5, 6); bit = 18; }
    else if (s == "Tetraloops") { rd.loops(tetra, 281, tetra_e, 6, 7); bit = 19; }
    else if (s == "Hexaloops") { rd.loops(hexa, 361, hexa_e, 8, 9); bit = 20; }
    else continue;  // unknown sections are reported and skipped by the reference
    if (!enth && bit >= 0) seen |= 1u << bit;
  }
  fclose(rd.fp);
  if (rd.failed) return fail(3, rd.err);
  if (seen != (1u << 21) - 1) return fail(4, "parameter file lacks a section the Zuker grammar reads");

  HostRna& H = *out;
  memset(&H, 0, sizeof H);
  DevRna& D = H.dev;
  H.lxc = kLxc37;
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b) D.stack[a][b] = as_short(stack[a][b]);
  for (int p = 0; p < 8; ++p)
    for (int x = 0; x < 5; ++x)
      for (int y = 0; y < 5; ++y) {
        D.mmH[p][x][y] = as_short(mm[0][p][x][y]);
        D.mmI[p][x][y] = as_short(mm[1][p][x][y]);
        D.mm1nI[p][x][y] = as_short(mm[2][p][x][y]);
        D.mm23I[p][x][y] = as_short(mm[3][p][x][y]);
        D.mmM[p][x][y] = as_short(clamp0(mm[4][p][x][y]));
        D.mmExt[p][x][y] = as_short(clamp0(mm[5][p][x][y]));
      }
  for (int p = 0; p < 8; ++p)
    for (int x = 0; x < 5; ++x) {
      D.d5[p][x] = as_short(clamp0(dng[0][p][x]));
      D.d3[p][x] = as_short(clamp0(dng[1][p][x]));
    }
  for (int l = 0; l <= kMaxLoop; ++l) {
    D.hairpin[l] = as_short(hp[l]);
    D.bulge[l] = as_short(bu[l]);
  }
  H.hairpin30 = (int16_t)D.hairpin[kMaxLoop];
  int16_t il30 = as_short(in[kMaxLoop]);
  for (int l = 0; l <= 60; ++l)
    D.il_ent[l] = l > kMaxLoop ? il30 + (int)(H.lxc * log((l) * (1.0 / kMaxLoop))) : as_short(in[l]);
  D.ninio0 = nin[0];
  D.ninio_max = nin[2];
  D.il5 = as_short(in[5]);
  D.ml_closing = as_short(mlp[2]);
  D.ml_intern = as_short(mlp[4]);
  D.term_au = as_short(misc[2]);
  for (int a = 0; a <= kMaxLoop; ++a)
    for (int b = 0; b <= kMaxLoop; ++b) {
      int asym = (a - b) * D.ninio0;
      if (asym < 0) asym = -asym;
      if (asym > D.ninio_max) asym = D.ninio_max;
      D.F[a][b] = (int16_t)(D.il_ent[a + b] + asym);
    }
  const char* strs[3] = {tri, tetra, hexa};
  const int* es[3] = {tri_e, tetra_e, hexa_e};
  const int steps[3] = {6, 7, 9}, lens[3] = {5, 6, 8};
  for (int c = 0; c < 3; ++c) {
    int L = (int)strlen(strs[c]);
    int cnt = 0;
    for (int off = 0; off < L && cnt < 40; off += steps[c], ++cnt) {
      D.loop_key[c][cnt] = loop_key(strs[c] + off, lens[c]);
      D.loop_e[c][cnt] = as_short(es[c][cnt]);
    }
    D.n_loop[c] = cnt;
  }
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b)
      for (int x = 0; x < 5; ++x)
        for (int y = 0; y < 5; ++y) {
          D.int11[a][b][x][y] = as_short(i11[a][b][x][y]);
          for (int z = 0; z < 5; ++z) {
            D.int21[a][b][x][y][z] = as_short(i21[a][b][x][y][z]);
            for (int w = 0; w < 5; ++w) D.int22[a][b][x][y][z][w] = as_short(i22[a][b][x][y][z][w]);
          }
        }
  return 0;
}

}  // namespace dpx
