/*
 * dpx_oracle.c -- CPU oracle (test infrastructure only; see dpx_oracle.h).
 *
 * Plain-C restatement of the reference algorithms, written from reading
 * them.  Citations are to files of manojo/lamp-dp-mt.
 */
#include "dpx_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>

/* energy_const.h:11,23-27 */
#define ORC_INF (INT_MAX / 10)
#define ORC_MAXLOOP 30
/* energy_data.h:53 -- lxc37 is the one constant the .par v2.0 file lacks */
static const double ORC_LXC37 = 107.856;

/* ======================================================================== */
/*  .par v2.0 reader (vienna.c:171-310)                                      */
/* ======================================================================== */

typedef struct {
  int stack[8][8];
  int mmH[8][5][5], mmI[8][5][5], mm1nI[8][5][5], mm23I[8][5][5], mmM[8][5][5], mmExt[8][5][5];
  int d5[8][5], d3[8][5];
  int int11[8][8][5][5], int21[8][8][5][5][5], int22[8][8][5][5][5][5];
  int hairpin[31], bulge[31], interior[31];
  int ml[6], ninio[3], misc[4];
  char tri[241], tetra[281], hexa[361];
  int triE[40], tetraE[40], hexaE[40];
  int dummy[40];
} raw_par_t;

/* get_array1 (vienna.c:232-272): fill arr[0..size) from successive lines. */
static int rd_values(FILE* fp, int* arr, int size) {
  int i = 0, last = 0;
  char line[8192];
  while (i < size) {
    if (!fgets(line, sizeof line, fp)) return -1;
    char* nl = strchr(line, '\n');
    if (nl) *nl = 0;
    char* c1 = strstr(line, "/*");
    if (c1) {  /* one single-line comment is removed (vienna.c:245-249) */
      char* c2 = strstr(c1, "*/");
      if (!c2) return -1;
      memmove(c1, c2 + 2, strlen(c2 + 2) + 1);
    }
    int pos = 0, pp = 0;
    char buf[16];
    while (i < size && sscanf(line + pos, "%15s%n", buf, &pp) == 1) {
      int p;
      pos += pp;
      if (buf[0] == '*') { i++; continue; }  /* keep previous value */
      if (buf[0] == 'x') {                   /* log extrapolation */
        if (i == 0) return -1;
        p = arr[last] + (int)(0.5 + ORC_LXC37 * log(((double)i) / (double)(last)));
      } else if (strcmp(buf, "DEF") == 0) p = -50;
      else if (strcmp(buf, "INF") == 0) p = ORC_INF;
      else if (strcmp(buf, "NST") == 0) p = 0;
      else {
        if (sscanf(buf, "%d", &p) != 1) return -1;
        last = i;
      }
      arr[i++] = p;
    }
  }
  return 0;
}

/* rd_Ndim_slice (vienna.c:129-158): recurse over the leading dimension from
 * pre[0] to dims[0]-post[0]; a slice with no shifts is one flat read. */
static int rd_slice(FILE* fp, int* arr, int nd, const int* dims, const int* pre, const int* post) {
  int shift = 0, inner = 1;
  for (int d = 0; d < nd; d++) shift += pre[d] + post[d];
  for (int d = 1; d < nd; d++) inner *= dims[d];
  if (shift == 0) return rd_values(fp, arr, inner * dims[0]);
  if (nd == 1) return rd_values(fp, arr + pre[0], dims[0] - pre[0] - post[0]);
  for (int i = pre[0]; i < dims[0] - post[0]; i++)
    if (rd_slice(fp, arr + i * inner, nd - 1, dims + 1, pre + 1, post + 1)) return -1;
  return 0;
}

/* rd_Tetraloop37 & co (vienna.c:276-310).  The line that ends the list is
 * consumed, and each iteration appends one blank, as the reference does. */
static void rd_loops(FILE* fp, char* str, int strsize, int* e37, int width, int stride) {
  char line[8192], fmt[32];
  int i = 0, r, dh;
  memset(str, 0, strsize);
  memset(e37, 0, sizeof(int) * 40);
  snprintf(fmt, sizeof fmt, "%%%ds %%d %%d", width);
  do {
    if (!fgets(line, sizeof line, fp)) break;
    r = sscanf(line, fmt, &str[stride * i], &e37[i], &dh);
    strcat(str, " ");
    i++;
  } while (r == 3 && i < 40);
}

enum { S_STACK, S_MMH, S_MMI, S_MM1N, S_MM23, S_MMM, S_MMEXT, S_D5, S_D3, S_INT11, S_INT21, S_INT22,
       S_HP, S_BU, S_IN, S_ML, S_NINIO, S_MISC, S_TRI, S_TETRA, S_HEXA, S_NSEC };

int orc_load_params(const char* path, orc_params_t* out) {
  FILE* fp = fopen(path, "r");
  if (!fp) return 1;
  static raw_par_t R; /* the reference keeps these arrays static too */
  static int junk[8 * 8 * 5 * 5 * 5 * 5];
  memset(&R, 0, sizeof R);
  int seen[S_NSEC] = {0};
  char line[8192], ident[256];
  if (!fgets(line, sizeof line, fp)) { fclose(fp); return 2; }
  if (strncmp(line, "## RNAfold parameter file v2.0", 30) != 0) { fclose(fp); return 3; }
  const int d2[2] = {8, 8}, s11[2] = {1, 1}, z2[2] = {0, 0};
  const int d3[3] = {8, 5, 5}, s100[3] = {1, 0, 0}, z3[3] = {0, 0, 0};
  const int d4[4] = {8, 8, 5, 5}, s4[4] = {1, 1, 0, 0}, z4[4] = {0, 0, 0, 0};
  const int d5[5] = {8, 8, 5, 5, 5}, s5[5] = {1, 1, 0, 0, 0}, z5[5] = {0, 0, 0, 0, 0};
  const int d6[6] = {8, 8, 5, 5, 5, 5}, s6[6] = {1, 1, 1, 1, 1, 1}, p6[6] = {1, 1, 0, 0, 0, 0};
  const int dd[2] = {8, 5}, sd[2] = {1, 0};
  int rc = 0;
  while (fgets(line, sizeof line, fp)) {
    if (sscanf(line, "# %255s", ident) != 1) continue;
#define ID(X) (strcmp(ident, X) == 0)
    int* tgt3 = NULL; int sec = -1;
    if (ID("stack")) { rc |= rd_slice(fp, &R.stack[0][0], 2, d2, s11, z2); seen[S_STACK] = 1; }
    else if (ID("mismatch_hairpin")) { tgt3 = &R.mmH[0][0][0]; sec = S_MMH; }
    else if (ID("mismatch_interior")) { tgt3 = &R.mmI[0][0][0]; sec = S_MMI; }
    else if (ID("mismatch_interior_1n")) { tgt3 = &R.mm1nI[0][0][0]; sec = S_MM1N; }
    else if (ID("mismatch_interior_23")) { tgt3 = &R.mm23I[0][0][0]; sec = S_MM23; }
    else if (ID("mismatch_multi")) { tgt3 = &R.mmM[0][0][0]; sec = S_MMM; }
    else if (ID("mismatch_exterior")) { tgt3 = &R.mmExt[0][0][0]; sec = S_MMEXT; }
    else if (ID("dangle5")) { rc |= rd_slice(fp, &R.d5[0][0], 2, dd, sd, z2); seen[S_D5] = 1; }
    else if (ID("dangle3")) { rc |= rd_slice(fp, &R.d3[0][0], 2, dd, sd, z2); seen[S_D3] = 1; }
    else if (ID("int11")) { rc |= rd_slice(fp, &R.int11[0][0][0][0], 4, d4, s4, z4); seen[S_INT11] = 1; }
    else if (ID("int21")) { rc |= rd_slice(fp, &R.int21[0][0][0][0][0], 5, d5, s5, z5); seen[S_INT21] = 1; }
    else if (ID("int22")) { rc |= rd_slice(fp, &R.int22[0][0][0][0][0][0], 6, d6, s6, p6); seen[S_INT22] = 1; }
    else if (ID("hairpin")) { rc |= rd_values(fp, R.hairpin, 31); seen[S_HP] = 1; }
    else if (ID("bulge")) { rc |= rd_values(fp, R.bulge, 31); seen[S_BU] = 1; }
    else if (ID("interior")) { rc |= rd_values(fp, R.interior, 31); seen[S_IN] = 1; }
    else if (ID("ML_params")) { rc |= rd_values(fp, R.ml, 6); seen[S_ML] = 1; }
    else if (ID("NINIO")) { rc |= rd_values(fp, R.ninio, 3); seen[S_NINIO] = 1; }
    else if (ID("Misc")) { rc |= rd_values(fp, R.misc, 4); seen[S_MISC] = 1; }
    else if (ID("Triloops")) { rd_loops(fp, R.tri, 241, R.triE, 5, 6); seen[S_TRI] = 1; }
    else if (ID("Tetraloops")) { rd_loops(fp, R.tetra, 281, R.tetraE, 6, 7); seen[S_TETRA] = 1; }
    else if (ID("Hexaloops")) { rd_loops(fp, R.hexa, 361, R.hexaE, 8, 9); seen[S_HEXA] = 1; }
    else if (ID("END")) break;
    else if (strstr(ident, "_enthalpies")) {
      /* enthalpies only matter away from 37 C; consume them with the same shape */
      const char* base = ident;
      if (!strncmp(base, "stack", 5)) rc |= rd_slice(fp, junk, 2, d2, s11, z2);
      else if (!strncmp(base, "mismatch", 8)) rc |= rd_slice(fp, junk, 3, d3, s100, z3);
      else if (!strncmp(base, "dangle", 6)) rc |= rd_slice(fp, junk, 2, dd, sd, z2);
      else if (!strncmp(base, "int11", 5)) rc |= rd_slice(fp, junk, 4, d4, s4, z4);
      else if (!strncmp(base, "int21", 5)) rc |= rd_slice(fp, junk, 5, d5, s5, z5);
      else if (!strncmp(base, "int22", 5)) rc |= rd_slice(fp, junk, 6, d6, s6, p6);
      else rc |= rd_values(fp, junk, 31);
    }
    if (tgt3) { rc |= rd_slice(fp, tgt3, 3, d3, s100, z3); seen[sec] = 1; }
#undef ID
    if (rc) break;
  }
  fclose(fp);
  if (rc) return 4;
  for (int s = 0; s < S_NSEC; s++) if (!seen[s]) return 5;

  /* get_scaled_parameters at 37 C (vienna.c:13-96): tempf == 1.0 exactly, so
   * RESCALE(dH,37) == the 37 C value; assignment to short truncates through
   * int as the x86-64 C conversion does. */
  orc_params_t* P = out;
  memset(P, 0, sizeof *P);
#define S16(v) ((int16_t)(int32_t)(v))
  for (int i = 0; i < 31; i++) {
    P->hairpin[i] = S16(R.hairpin[i]);
    P->bulge[i] = S16(R.bulge[i]);
    P->internal_loop[i] = S16(R.interior[i]);
  }
  P->lxc = ORC_LXC37;
  P->ninio[0] = R.ninio[0];
  P->ninio[1] = R.ninio[2];
  P->TerminalAU = S16(R.misc[2]);
  P->DuplexInit = S16(R.misc[0]);
  P->MLbase = S16(R.ml[0]);
  for (int i = 0; i < 8; i++) P->MLintern[i] = S16(R.ml[4]);
  P->MLclosing = S16(R.ml[2]);
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++) P->stack[i][j] = S16(R.stack[i][j]);
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 5; j++)
      for (int k = 0; k < 5; k++) {
        P->mismatchI[i][j][k] = S16(R.mmI[i][j][k]);
        P->mismatchH[i][j][k] = S16(R.mmH[i][j][k]);
        P->mismatch1nI[i][j][k] = S16(R.mm1nI[i][j][k]);
        P->mismatch23I[i][j][k] = S16(R.mm23I[i][j][k]);
        int mm = R.mmM[i][j][k]; P->mismatchM[i][j][k] = S16(mm > 0 ? 0 : mm);
        mm = R.mmExt[i][j][k]; P->mismatchExt[i][j][k] = S16(mm > 0 ? 0 : mm);
      }
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 5; j++) {
      int dd5 = R.d5[i][j], dd3 = R.d3[i][j];
      P->dangle5[i][j] = S16(dd5 > 0 ? 0 : dd5);
      P->dangle3[i][j] = S16(dd3 > 0 ? 0 : dd3);
    }
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++)
      for (int k = 0; k < 5; k++)
        for (int l = 0; l < 5; l++) {
          P->int11[i][j][k][l] = S16(R.int11[i][j][k][l]);
          for (int m = 0; m < 5; m++) {
            P->int21[i][j][k][l][m] = S16(R.int21[i][j][k][l][m]);
            for (int q = 0; q < 5; q++) P->int22[i][j][k][l][m][q] = S16(R.int22[i][j][k][l][m][q]);
          }
        }
  P->TetraloopsLen = (int16_t)strlen(R.tetra);
  P->TriloopsLen = (int16_t)strlen(R.tri);
  P->HexaloopsLen = (int16_t)strlen(R.hexa);
  for (int i = 0; i * 7 < P->TetraloopsLen && i < 40; i++) P->Tetraloop_E[i] = S16(R.tetraE[i]);
  for (int i = 0; i * 5 < P->TriloopsLen && i < 40; i++) P->Triloop_E[i] = S16(R.triE[i]);
  for (int i = 0; i * 9 < P->HexaloopsLen && i < 40; i++) P->Hexaloop_E[i] = S16(R.hexaE[i]);
  strncpy(P->Tetraloops, R.tetra, 281);
  strncpy(P->Triloops, R.tri, 241);
  strncpy(P->Hexaloops, R.hexa, 361);
#undef S16
  return 0;
}

/* ======================================================================== */
/*  librna energy model (src/librna/librna_impl.h:106-293), GAPS_DISABLED,   */
/*  NOCHECKS, codes N=0 A=1 C=2 G=3 U=4                                      */
/* ======================================================================== */

/* bp_index (librna_impl.h:106-114): CG=1 GC=2 GU=3 UG=4 AU=5 UA=6 none=7 */
int orc_bp_index(int x, int y) {
  switch (x) {
    case 1: if (y == 4) return 5; break;
    case 2: if (y == 3) return 1; break;
    case 3: if (y == 2) return 2; if (y == 4) return 3; break;
    case 4: if (y == 3) return 4; if (y == 1) return 6; break;
  }
  return 7;
}
#define SQ(k) ((int)r->s[k])
#define BP(a, b) orc_bp_index(SQ(a), SQ(b))

/* jacobson_stockmayer (librna_impl.h:149) */
static int orc_js(const orc_rna_t* r, int l) { return (int)(r->P->lxc * log((l) * (1.0 / ORC_MAXLOOP))); }
static int orc_hl_ent(const orc_rna_t* r, int l) { return l > ORC_MAXLOOP ? r->P->hairpin[ORC_MAXLOOP] + orc_js(r, l) : r->P->hairpin[l]; }
static int orc_hl_stack(const orc_rna_t* r, int i, int j) { return r->P->mismatchH[BP(i, j)][SQ(i + 1)][SQ(j - 1)]; }

/* termau_energy (librna_impl.h:152) */
int orc_termau_energy(const orc_rna_t* r, int i, int j) {
  return ((SQ(i) == 3 && SQ(j) == 2) || (SQ(i) == 2 && SQ(j) == 3)) ? 0 : r->P->TerminalAU;
}

/* strMatch (librna_impl.h:154-179): index of the first entry of str (entries
 * every `step` chars) equal to the bases in[i..j]. */
static int orc_str_match(const orc_rna_t* r, int i, int j, const char* str, int strlen_, int step) {
  static const char chr[] = "NACGU";
  int len = j - i + 1;
  for (const char* p = str; p < str + strlen_; p += step) {
    int m;
    for (m = 0; m < len; ++m) if (p[m] != chr[SQ(i + m)]) break;
    if (m == len) return (int)(p - str) / step;
  }
  return -1;
}

/* hl_energy (librna_impl.h:181-193) */
int orc_hl_energy(const orc_rna_t* r, int i, int j) {
  const orc_params_t* P = r->P;
  int size = j - i - 1, pos;
  if (size < 3) return 600;
  switch (size) {
    case 3: if ((pos = orc_str_match(r, i, j, P->Triloops, P->TriloopsLen, 6)) != -1) return P->Triloop_E[pos]; break;
    case 4: if ((pos = orc_str_match(r, i, j, P->Tetraloops, P->TetraloopsLen, 7)) != -1) return P->Tetraloop_E[pos]; break;
    case 6: if ((pos = orc_str_match(r, i, j, P->Hexaloops, P->HexaloopsLen, 9)) != -1) return P->Hexaloop_E[pos]; break;
  }
  int ent = orc_hl_ent(r, size);
  if (size == 3) return ent + orc_termau_energy(r, i, j);
  return ent + orc_hl_stack(r, i, j);
}

static int orc_bl_ent(const orc_rna_t* r, int l) { return l > ORC_MAXLOOP ? r->P->bulge[ORC_MAXLOOP] + orc_js(r, l) : r->P->bulge[l]; }

/* bl_energy (librna_impl.h:199-206); getPrev(j,1,x) == j-1 under NOCHECKS */
int orc_bl_energy(const orc_rna_t* r, int i, int k, int l, int j, int xr) {
  (void)xr;
  int size = l - k + 1;
  if (size == 0) return r->P->stack[BP(i, j)][BP(j - 1, l + 1)];
  if (size == 1) return orc_bl_ent(r, size) + r->P->stack[BP(i, j)][BP(j - 1, l + 1)];
  if (size > 1) return orc_bl_ent(r, size) + orc_termau_energy(r, i, j) + orc_termau_energy(r, j - 1, l + 1);
  return -1000000;
}

/* br_energy (librna_impl.h:208-215); getNext(i,1,x) == i+1 */
int orc_br_energy(const orc_rna_t* r, int i, int k, int l, int j, int xl) {
  (void)xl;
  int size = l - k + 1;
  if (size == 0) return r->P->stack[BP(i, j)][BP(k - 1, i + 1)];
  if (size == 1) return orc_bl_ent(r, size) + r->P->stack[BP(i, j)][BP(k - 1, i + 1)];
  if (size > 1) return orc_bl_ent(r, size) + orc_termau_energy(r, i, j) + orc_termau_energy(r, k - 1, i + 1);
  return -1000000;
}

static int orc_il_ent(const orc_rna_t* r, int l) { return l > ORC_MAXLOOP ? r->P->internal_loop[ORC_MAXLOOP] + orc_js(r, l) : r->P->internal_loop[l]; }
static int orc_il_asym(const orc_rna_t* r, int sl, int sr) {
  int v = (sl - sr) * r->P->ninio[0];
  if (v < 0) v = -v;
  return v < r->P->ninio[1] ? v : r->P->ninio[1];
}

/* il_energy (librna_impl.h:230-265) with the il11/12/21/22 lookups (:218-221) */
int orc_il_energy(const orc_rna_t* r, int i, int k, int l, int j) {
  const orc_params_t* P = r->P;
  int sl = k - i - 1, sr = j - l - 1;
  if (sl == 0) return orc_br_energy(r, i, l + 1, j - 1, j, k);
  if (sr == 0) return orc_bl_energy(r, i, i + 1, k - 1, j, l);
  int ocb = BP(i, j), olb = SQ(i + 1), orb = SQ(j - 1);
  int icb = BP(l, k), ilb = SQ(l + 1), irb = SQ(k - 1);
  if (sl == 1) {
    if (sr == 1) return P->int11[ocb][BP(j - 2, i + 2)][SQ(i + 1)][SQ(j - 1)];
    if (sr == 2) return P->int21[ocb][BP(j - 3, i + 2)][SQ(i + 1)][SQ(j - 2)][SQ(j - 1)];
    return orc_il_ent(r, sl + sr) + orc_il_asym(r, sl, sr) + P->mismatch1nI[ocb][olb][orb] + P->mismatch1nI[icb][ilb][irb];
  } else if (sl == 2) {
    if (sr == 1) return P->int21[BP(j - 2, i + 3)][ocb][SQ(j - 1)][SQ(i + 1)][SQ(i + 2)];
    if (sr == 2) return P->int22[ocb][BP(j - 3, i + 3)][SQ(i + 1)][SQ(i + 2)][SQ(j - 2)][SQ(j - 1)];
    if (sr == 3) return P->internal_loop[5] + P->ninio[0] + P->mismatch23I[ocb][olb][orb] + P->mismatch23I[icb][ilb][irb];
  } else if (sl == 3 && sr == 2) {
    return P->internal_loop[5] + P->ninio[0] + P->mismatch23I[ocb][olb][orb] + P->mismatch23I[icb][ilb][irb];
  } else if (sr == 1) {
    return orc_il_ent(r, sl + sr) + orc_il_asym(r, sl, sr) + P->mismatch1nI[ocb][olb][orb] + P->mismatch1nI[icb][ilb][irb];
  }
  return orc_il_ent(r, sl + sr) + P->mismatchI[ocb][olb][orb] + P->mismatchI[icb][ilb][irb] + orc_il_asym(r, sl, sr);
}

/* sr_energy (librna_impl.h:267) */
int orc_sr_energy(const orc_rna_t* r, int i, int j) { return r->P->stack[BP(i, j)][BP(j - 1, i + 1)]; }

/* ext_mismatch_energy (librna_impl.h:282-287) with NOCHECKS dl/dr (:271-272) */
int orc_ext_mismatch_energy(const orc_rna_t* r, int i, int j) {
  const orc_params_t* P = r->P;
  if (i > 0 && j + 1 < r->n) return P->mismatchExt[BP(i, j)][SQ(i - 1)][SQ(j + 1)];
  if (i > 0) return P->dangle5[BP(i, j)][SQ(i - 1)];
  if (j + 1 < r->n) return P->dangle3[BP(i, j)][SQ(j + 1)];
  return 0;
}

/* ml_mismatch_energy (librna_impl.h:289) */
int orc_ml_mismatch_energy(const orc_rna_t* r, int i, int j) {
  return r->P->mismatchM[BP(j, i)][SQ(j - 1)][SQ(i + 1)];
}

/* ======================================================================== */
/*  Zuker (bench2/zu_kern.h:92-184 BACKTRACK branch; zu.c:23-111)           */
/* ======================================================================== */

#define ZU_EMPTY 999999
typedef struct { int cl, ml, ml1, st; } zu_cost_t;
typedef struct { int rule, p0, p1; } zu_bt_t;
typedef struct { zu_bt_t cl, ml, ml1, st; } zu_back_t;

int32_t orc_zuker(const orc_params_t* P, const uint8_t* seq, int n,
                  orc_trace_t* trace, int trace_cap, int* trace_len, int32_t* tables) {
  if (n < 0 || !P) return INT32_MIN;
  for (int q = 0; q < n; q++) if (seq[q] < 1 || seq[q] > 4) return INT32_MIN;
  const int W = n + 1, H = n + 1;
  orc_rna_t R = {P, seq, n};
  const orc_rna_t* r = &R;
  zu_cost_t* cost = (zu_cost_t*)calloc((size_t)W * H, sizeof(zu_cost_t));
  zu_back_t* back = (zu_back_t*)calloc((size_t)W * H, sizeof(zu_back_t));
  if (!cost || !back) { free(cost); free(back); return INT32_MIN; }
#define IX(a, b) ((size_t)(a) * W + (b))
#define VALID(a, b, f) (cost[IX(a, b)].f != ZU_EMPTY)
#define TAKE(val, rl, q0, q1) do { int _c = (val); if (_c < c0) { c0 = _c; b0.rule = (rl); b0.p0 = (q0); b0.p1 = (q1); } } while (0)
  /* fun8/fun9 (zu.c:37,38): dangle-closed helix entering a multiloop or the exterior */
#define DLR(i_, k_) (cost[IX(i_, k_)].cl + orc_ext_mismatch_energy(r, (i_), (k_) - 1) + orc_termau_energy(r, (i_), (k_) - 1))
  const int mlint = P->MLintern[0], mlclose = P->MLclosing;
  for (int jj = 0; jj < W; ++jj) {
    for (int ii = jj; ii < H; ++ii) {
      const int i = H - 1 - ii, j = i + jj;
      { /* ---- cl[i,j] ---- */
        int c0 = ZU_EMPTY; zu_bt_t b0 = {-1, 0, 0};
        if (i + 4 <= j && BP(i, j - 1) != 7 && BP(i + 1, j - 2) != 7) { /* fun6 */
          const int sr = orc_sr_energy(r, i, j - 1);
          if (i + 9 <= j && VALID(i + 1, j - 1, cl)) TAKE(cost[IX(i + 1, j - 1)].cl + sr, 0, 0, 0);
          if (i + 7 <= j) TAKE(orc_hl_energy(r, i + 1, j - 2) + sr, 1, 0, 0);
          if (i + 12 <= j) {
            int ku = j - 9 < i + 32 ? j - 9 : i + 32;
            for (int k = i + 3; k <= ku; ++k)
              if (VALID(k, j - 2, cl)) TAKE(cost[IX(k, j - 2)].cl + orc_bl_energy(r, i + 1, i + 2, k - 1, j - 2, j - 3) + sr, 2, k, 0);
          }
          if (i + 12 <= j) {
            int kl = i + 9 > j - 32 ? i + 9 : j - 32;
            for (int k = kl; k <= j - 3; ++k)
              if (VALID(i + 2, k, cl)) TAKE(cost[IX(i + 2, k)].cl + orc_br_energy(r, i + 1, k, j - 3, j - 2, i + 2) + sr, 3, k, 0);
          }
          if (i + 13 <= j) {
            int kl = i + 10 > j - 32 ? i + 10 : j - 32;
            for (int k = kl; k <= j - 3; ++k) {
              int lu = k - 7 < i + 32 ? k - 7 : i + 32;
              for (int l = i + 3; l <= lu; ++l)
                if (VALID(l, k, cl)) TAKE(cost[IX(l, k)].cl + orc_il_energy(r, i + 1, l, k - 1, j - 2) + sr, 4, l, k);
            }
          }
          if (i + 18 <= j && VALID(i + 2, j - 2, ml))
            TAKE(mlclose + mlint + cost[IX(i + 2, j - 2)].ml + orc_termau_energy(r, i + 1, j - 2) + sr +
                     orc_ml_mismatch_energy(r, i + 1, j - 2), 5, 0, 0);
        }
        cost[IX(i, j)].cl = c0; back[IX(i, j)].cl = b0;
      }
      { /* ---- ml[i,j] ---- */
        int c0 = ZU_EMPTY; zu_bt_t b0 = {-1, 0, 0};
        if (i + 15 <= j && VALID(i + 1, j, ml)) TAKE(cost[IX(i + 1, j)].ml, 6, 0, 0);
        for (int k = i + 7; k <= j - 7; ++k)
          if (VALID(i, k, cl) && VALID(k, j, ml1)) TAKE(mlint + DLR(i, k) + cost[IX(k, j)].ml1, 7, k, 0);
        cost[IX(i, j)].ml = c0; back[IX(i, j)].ml = b0;
      }
      { /* ---- ml1[i,j] ---- */
        int c0 = ZU_EMPTY; zu_bt_t b0 = {-1, 0, 0};
        if (i + 8 <= j && VALID(i + 1, j, ml1)) TAKE(cost[IX(i + 1, j)].ml1, 8, 0, 0);
        for (int k = i + 7; k <= j - 7; ++k)
          if (VALID(i, k, cl) && VALID(k, j, ml1)) TAKE(mlint + DLR(i, k) + cost[IX(k, j)].ml1, 9, k, 0);
        if (i + 7 <= j && VALID(i, j, cl)) TAKE(mlint + DLR(i, j), 10, 0, 0);
        for (int k = i + 7; k < j; ++k)
          if (VALID(i, k, cl)) TAKE(mlint + DLR(i, k), 11, k, 0);
        cost[IX(i, j)].ml1 = c0; back[IX(i, j)].ml1 = b0;
      }
      { /* ---- st[i,j] ---- */
        int c0 = ZU_EMPTY; zu_bt_t b0 = {-1, 0, 0};
        if (j == W - 1) { /* fun14 */
          if (i + 1 <= j) TAKE(cost[IX(i + 1, j)].st, 12, 0, 0);
          for (int k = i + 7; k <= j; ++k)
            if (VALID(i, k, cl)) TAKE(DLR(i, k) + cost[IX(k, j)].st, 13, k, 0);
          if (i == j) TAKE(0, 14, 0, 0);
        }
        cost[IX(i, j)].st = c0; back[IX(i, j)].st = b0;
      }
    }
  }
  const int32_t res = cost[IX(0, W - 1)].st;
  if (tables) {
    for (size_t q = 0; q < (size_t)W * H; q++) {
      tables[4 * q + 0] = cost[q].cl; tables[4 * q + 1] = cost[q].ml;
      tables[4 * q + 2] = cost[q].ml1; tables[4 * q + 3] = cost[q].st;
    }
  }
  if (trace && trace_len) {
    /* cpu_backtrack (zu.c:49-93): BFS from (0,n) rule 12 */
    static const int tl[15] = {2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1};
    int rd = 0, wr = 0;
#define PUSH(a, b, rl) do { if (wr < trace_cap) { trace[wr].i = (a); trace[wr].j = (b); trace[wr].rule = (rl); trace[wr].pos[0] = trace[wr].pos[1] = 0; } wr++; } while (0)
#define MX(a, b) ((a) > (b) ? (a) : (b))
    PUSH(0, W - 1, 12);
    for (; rd < wr && rd < trace_cap; ++rd) {
      orc_trace_t* t = &trace[rd];
      const zu_back_t* bk = &back[IX(t->i, t->j)];
      const zu_bt_t* bt = t->rule <= 5 ? &bk->cl : t->rule <= 7 ? &bk->ml : t->rule <= 11 ? &bk->ml1 : &bk->st;
      t->rule = bt->rule;
      if (t->rule < 0) break;
      t->pos[0] = bt->p0;
      if (tl[t->rule] > 1) t->pos[1] = bt->p1;
      const int i = t->i, j = t->j, p = t->pos[0], q = t->pos[1];
      switch (t->rule) {
        case 0: PUSH(i + 1, MX(i + 8, j - 1), 0); break;
        case 1: break;
        case 2: PUSH(p, MX(i + 10, MX(i + 11, j - 1) - 1), 0); break;
        case 3: PUSH(i + 2, p, 0); break;
        case 4: PUSH(p, q, 0); break;
        case 5: PUSH(i + 2, MX(i + 16, MX(i + 17, j - 1) - 1), 6); break;
        case 6: PUSH(i + 1, j, 6); break;
        case 7: PUSH(i, MX(i + 7, p), 0); PUSH(p, j, 8); break;
        case 8: PUSH(i + 1, j, 8); break;
        case 9: PUSH(i, MX(i + 7, p), 0); PUSH(p, j, 8); break;
        case 10: PUSH(i, MX(i + 7, j), 0); break;
        case 11: PUSH(i, MX(i + 7, p), 0); break;
        case 12: PUSH(i + 1, j, 12); break;
        case 13: PUSH(i, MX(i + 7, p), 0); PUSH(p, j, 12); break;
        case 14: break;
      }
    }
    *trace_len = wr;
#undef PUSH
  }
  free(cost); free(back);
  return res;
#undef TAKE
#undef VALID
#undef DLR
}

/* ======================================================================== */
/*  Nussinov (bench2/nu_kern.h:23-45; nu.c:17-73; Nussinov.scala:49-76)     */
/* ======================================================================== */
/* RNAfold grammar (RNAFold.scala:20-64); restated from the grammar and the
 * C code generator's conventions (CodeGen.scala:100-208: alternatives in
 * order, split indices ascending with the outer split outside, a candidate
 * replaces the best only when strictly smaller, the first always). */

typedef struct { int32_t c; int32_t rule, p0, p1; } rf_slot_t;

int32_t orc_rnafold(const orc_params_t* P, const uint8_t* seq, int n,
                    orc_trace_t* trace, int trace_cap, int* trace_len, int32_t* tables) {
  if (n < 0 || !P) return INT32_MIN;
  for (int q = 0; q < n; q++) if (seq[q] < 1 || seq[q] > 4) return INT32_MIN;
  const int W = n + 1;
  orc_rna_t R = {P, seq, n};
  const orc_rna_t* r = &R;
  rf_slot_t* Q = (rf_slot_t*)malloc(sizeof(rf_slot_t) * (size_t)W * W);
  rf_slot_t* QM = (rf_slot_t*)malloc(sizeof(rf_slot_t) * (size_t)W * W);
  rf_slot_t* QP = (rf_slot_t*)malloc(sizeof(rf_slot_t) * (size_t)W * W);
  if (!Q || !QM || !QP) { free(Q); free(QM); free(QP); return INT32_MIN; }
  for (size_t q = 0; q < (size_t)W * W; q++) {
    Q[q].c = QM[q].c = QP[q].c = RF_EMPTY;
    Q[q].rule = QM[q].rule = QP[q].rule = -1;
  }
#define RX(a, b) ((size_t)(a) * W + (b))
#define RTAKE(slot, val, rl, q0, q1) do { int _c = (val); if ((slot).rule == -1 || _c < (slot).c) { \
    (slot).c = _c; (slot).rule = (rl); (slot).p0 = (q0); (slot).p1 = (q1); } } while (0)
  for (int d = 0; d <= n; ++d) {
    for (int i = 0; i + d <= n; ++i) {
      const int j = i + d;
      /* Qp: basepairing (ADPParsers.scala:96) */
      if (i + 2 <= j && BP(i, j - 1) != 7) {
        rf_slot_t* s = &QP[RX(i, j)];
        if (j - i >= 3) RTAKE(*s, orc_hl_energy(r, i, j - 1), 5, 0, 0);
        if (QP[RX(i + 1, j - 1)].rule != -1 && i + 1 <= j - 1)
          RTAKE(*s, orc_sr_energy(r, i, j - 1) + QP[RX(i + 1, j - 1)].c, 6, 0, 0);
        {
          const int l0 = i + 4 > j - 30 ? i + 4 : j - 30;
          for (int l = l0; l <= j - 1; ++l) {
            const int k1 = i + 30 < l - 3 ? i + 30 : l - 3;
            for (int k = i + 1; k <= k1; ++k)
              if (QP[RX(k, l)].rule != -1) RTAKE(*s, orc_il_energy(r, i, k, l - 1, j - 1) + QP[RX(k, l)].c, 7, k, l);
          }
        }
        if (i + 1 <= j - 1 && QM[RX(i + 1, j - 1)].rule != -1) RTAKE(*s, QM[RX(i + 1, j - 1)].c, 8, 0, 0);
      }
      /* QM: Q ~ Q, filter length(4) */
      if (j - i >= 4) {
        rf_slot_t* s = &QM[RX(i, j)];
        for (int k = i + 3; k <= j - 3; ++k)
          if (Q[RX(i, k)].rule != -1 && Q[RX(k, j)].rule != -1) RTAKE(*s, Q[RX(i, k)].c + Q[RX(k, j)].c, 4, k, 0);
      }
      /* Q: filter length(2) */
      if (j - i >= 2) {
        rf_slot_t* s = &Q[RX(i, j)];
        if (QM[RX(i, j)].rule != -1) RTAKE(*s, QM[RX(i, j)].c, 0, 0, 0);
        if (Q[RX(i, j - 1)].rule != -1) RTAKE(*s, Q[RX(i, j - 1)].c, 1, 0, 0);
        if (Q[RX(i + 1, j)].rule != -1) RTAKE(*s, Q[RX(i + 1, j)].c, 2, 0, 0);
        if (QP[RX(i, j)].rule != -1) RTAKE(*s, QP[RX(i, j)].c, 3, 0, 0);
      }
    }
  }
  const int32_t res = Q[RX(0, n)].rule == -1 ? RF_EMPTY : Q[RX(0, n)].c;
  if (tables) {
    for (size_t q = 0; q < (size_t)W * W; q++) {
      tables[3 * q + 0] = Q[q].rule == -1 ? RF_EMPTY : Q[q].c;
      tables[3 * q + 1] = QM[q].rule == -1 ? RF_EMPTY : QM[q].c;
      tables[3 * q + 2] = QP[q].rule == -1 ? RF_EMPTY : QP[q].c;
    }
  }
  if (trace && trace_len) {
    int rd = 0, wr = 0;
#define RPUSH(a, b, rl) do { if (wr < trace_cap) { trace[wr].i = (a); trace[wr].j = (b); trace[wr].rule = (rl); trace[wr].pos[0] = trace[wr].pos[1] = 0; } wr++; } while (0)
    if (res != RF_EMPTY) RPUSH(0, n, 0);
    for (; rd < wr && rd < trace_cap; ++rd) {
      orc_trace_t* t = &trace[rd];
      const rf_slot_t* s = t->rule <= 3 ? &Q[RX(t->i, t->j)] : t->rule == 4 ? &QM[RX(t->i, t->j)] : &QP[RX(t->i, t->j)];
      const int i = t->i, j = t->j;
      t->rule = s->rule;
      if (t->rule == 4) t->pos[0] = s->p0;
      if (t->rule == 7) { t->pos[0] = s->p0; t->pos[1] = s->p1; }
      switch (t->rule) {
        case 0: RPUSH(i, j, 4); break;
        case 1: RPUSH(i, j - 1, 0); break;
        case 2: RPUSH(i + 1, j, 0); break;
        case 3: RPUSH(i, j, 5); break;
        case 4: RPUSH(i, s->p0, 0); RPUSH(s->p0, j, 0); break;
        case 6: RPUSH(i + 1, j - 1, 5); break;
        case 7: RPUSH(s->p0, s->p1, 5); break;
        case 8: RPUSH(i + 1, j - 1, 4); break;
        default: break;
      }
    }
    *trace_len = wr;
#undef RPUSH
  }
#undef RX
#undef RTAKE
  free(Q); free(QM); free(QP);
  return res;
}

/* ======================================================================== */

int32_t orc_nussinov(const uint8_t* seq, int n, int min_loop,
                     orc_trace_t* trace, int trace_cap, int* trace_len, int32_t* table) {
  if (n < 0 || min_loop < 0) return INT32_MIN;
  for (int q = 0; q < n; q++) if (seq[q] < 1 || seq[q] > 4) return INT32_MIN;
  const int W = n + 1, H = n + 1;
  int* s = (int*)calloc((size_t)W * H, sizeof(int));
  int* br = (int*)calloc((size_t)W * H, sizeof(int));
  int* bp = (int*)calloc((size_t)W * H, sizeof(int));
  if (!s || !br || !bp) { free(s); free(br); free(bp); return INT32_MIN; }
  for (int jj = 0; jj < W; ++jj)
    for (int ii = jj; ii < H; ++ii) {
      const int i = H - 1 - ii, j = i + jj;
      int c0 = -1000, rule = -1, pk = 0;
#define TK(v, rl, k_) do { int _c = (v); if (_c > c0) { c0 = _c; rule = (rl); pk = (k_); } } while (0)
      if (i == j) TK(0, 0, 0);
      if (i + 1 <= j) TK(s[IX(i + 1, j)], 1, 0);
      if (i + 1 <= j) TK(s[IX(i, j - 1)], 2, 0);
      /* fun4 (nu.c:24): canonical pair in[i]..in[j-1]; min_loop generalises i+2<=j */
      if (i + 2 + min_loop <= j) {
        int a = seq[i], b = seq[j - 1];
        int ok = (a == 1 && b == 4) || (a == 4 && b == 1) || (a == 3 && b == 4) || (a == 4 && b == 3) ||
                 (a == 2 && b == 3) || (a == 3 && b == 2);
        if (ok) TK(s[IX(i + 1, j - 1)] + 1, 3, 0);
      }
      for (int k = i + 1; k < j; ++k) TK(s[IX(i, k)] + s[IX(k, j)], 4, k);
#undef TK
      s[IX(i, j)] = c0; br[IX(i, j)] = rule; bp[IX(i, j)] = pk;
    }
  const int32_t res = s[IX(0, W - 1)];
  if (table) memcpy(table, s, sizeof(int) * (size_t)W * H);
  if (trace && trace_len) {
    int rd = 0, wr = 0;
#define PUSH(a, b) do { if (wr < trace_cap) { trace[wr].i = (a); trace[wr].j = (b); trace[wr].rule = 0; trace[wr].pos[0] = trace[wr].pos[1] = 0; } wr++; } while (0)
    PUSH(0, W - 1);
    for (; rd < wr && rd < trace_cap; ++rd) {
      orc_trace_t* t = &trace[rd];
      t->rule = br[IX(t->i, t->j)];
      t->pos[0] = bp[IX(t->i, t->j)];
      const int i = t->i, j = t->j;
      switch (t->rule) { /* nu.c:50-56 */
        case 1: PUSH(i + 1, j); break;
        case 2: PUSH(i, MX(i, j - 1)); break;
        case 3: PUSH(i + 1, MX(i + 1, j - 1)); break;
        case 4: PUSH(i, t->pos[0]); PUSH(t->pos[0], j); break;
        default: break;
      }
    }
    *trace_len = wr;
#undef PUSH
  }
  free(s); free(br); free(bp);
  return res;
}

/* ======================================================================== */
/*  SW / Gotoh, two-track (bench/dp_swat/comp1.cu:50-132)                    */
/*  Rule ids: M 0..3 (empty, pair, g1, g2), g1 4..6 (zero1, open1, ext1),    */
/*  g2 7..9 (zero2, open2, ext2) -- comp1.h:2-4                              */
/* ======================================================================== */

int32_t orc_swat(const uint8_t* a, int n1, const uint8_t* b, int n2,
                 int match, int mismatch, int gap_open, int gap_ext,
                 int32_t* best, int* bi, int* bj, int from_best,
                 orc_trace_t* trace, int trace_cap, int* trace_len) {
  if (n1 < 0 || n2 < 0) return INT32_MIN;
  const size_t W = (size_t)n2 + 1, H = (size_t)n1 + 1;
  int* M = (int*)malloc(sizeof(int) * W * H);
  int* G1 = (int*)malloc(sizeof(int) * W * H);
  int* G2 = (int*)malloc(sizeof(int) * W * H);
  uint8_t* rm = (uint8_t*)malloc(W * H);
  uint8_t* r1 = (uint8_t*)malloc(W * H);
  uint8_t* r2 = (uint8_t*)malloc(W * H);
  if (!M || !G1 || !G2 || !rm || !r1 || !r2) { free(M); free(G1); free(G2); free(rm); free(r1); free(r2); return INT32_MIN; }
#define SX(i, j) ((size_t)(i) * W + (j))
  /* open/ext are max(0, a+cost) (SWatAffine.scala:26-29; comp1.cu fun1..fun4) */
#define CLAMP(v) ((v) > 0 ? (v) : 0)
  int32_t bestv = INT32_MIN; int besti = 0, bestj = 0;
  for (int i = 0; i <= n1; ++i)
    for (int j = 0; j <= n2; ++j) {
      int c, cv, rl;
      /* g2: the first applicable candidate is always taken, later ones on strict > */
      cv = 0; rl = -1;
      if (j == 0) { cv = 0; rl = 7; }
      if (j >= 1) { c = CLAMP(M[SX(i, j - 1)] + gap_open); if (c > cv || rl == -1) { cv = c; rl = 8; } }
      if (j >= 1) { c = CLAMP(G2[SX(i, j - 1)] + gap_ext); if (c > cv || rl == -1) { cv = c; rl = 9; } }
      G2[SX(i, j)] = cv; r2[SX(i, j)] = (uint8_t)rl;
      cv = 0; rl = -1;
      if (i == 0) { cv = 0; rl = 4; }
      if (i >= 1) { c = CLAMP(M[SX(i - 1, j)] + gap_open); if (c > cv || rl == -1) { cv = c; rl = 5; } }
      if (i >= 1) { c = CLAMP(G1[SX(i - 1, j)] + gap_ext); if (c > cv || rl == -1) { cv = c; rl = 6; } }
      G1[SX(i, j)] = cv; r1[SX(i, j)] = (uint8_t)rl;
      cv = 0; rl = -1;
      if (i == 0 && j == 0) { cv = 0; rl = 0; }
      if (i >= 1 && j >= 1) { c = M[SX(i - 1, j - 1)] + (a[i - 1] == b[j - 1] ? match : mismatch); if (c > cv || rl == -1) { cv = c; rl = 1; } }
      c = G1[SX(i, j)]; if (c > cv || rl == -1) { cv = c; rl = 2; }
      c = G2[SX(i, j)]; if (c > cv || rl == -1) { cv = c; rl = 3; }
      M[SX(i, j)] = cv; rm[SX(i, j)] = (uint8_t)rl;
      if (cv > bestv) { bestv = cv; besti = i; bestj = j; }
    }
  const int32_t res = M[SX(n1, n2)];
  if (best) *best = bestv;
  if (bi) *bi = besti;
  if (bj) *bj = bestj;
  if (trace && trace_len) {
    int rd = 0, wr = 0;
#define PUSH(x, y, rl_) do { if (wr < trace_cap) { trace[wr].i = (x); trace[wr].j = (y); trace[wr].rule = (rl_); trace[wr].pos[0] = trace[wr].pos[1] = 0; } wr++; } while (0)
    if (from_best) PUSH(besti, bestj, 0); else PUSH(n1, n2, 0);
    for (; rd < wr && rd < trace_cap; ++rd) {
      orc_trace_t* t = &trace[rd];
      const int i = t->i, j = t->j;
      t->rule = t->rule <= 3 ? rm[SX(i, j)] : t->rule <= 6 ? r1[SX(i, j)] : r2[SX(i, j)];
      switch (t->rule) { /* comp1.cu:121-131 */
        case 1: PUSH(i - 1, j - 1, 0); break;
        case 2: PUSH(i, j, 4); break;
        case 3: PUSH(i, j, 7); break;
        case 5: PUSH(i - 1, j, 0); break;
        case 6: PUSH(i - 1, j, 4); break;
        case 8: PUSH(i, j - 1, 0); break;
        case 9: PUSH(i, j - 1, 7); break;
        default: break;
      }
    }
    *trace_len = wr;
#undef PUSH
  }
  free(M); free(G1); free(G2); free(rm); free(r1); free(r2);
  return res;
#undef SX
#undef CLAMP
}

/* ======================================================================== */
/*  Matrix chain, Int algebra (Benchmarks.scala:18-27; MatrixMult2.scala)  */
/*  rule 0 = single (el, j==i+1), rule 1 = mult over k in [i+1,j-1];        */
/*  min, first candidate taken then strict < (CodeGen.scala:122-126)         */
/* ======================================================================== */

int32_t orc_matchain(const int32_t* rows, const int32_t* cols, int n,
                     orc_trace_t* trace, int trace_cap, int* trace_len) {
  if (n < 1) return INT32_MIN;
  const int W = n + 1, H = n + 1;
  int32_t* m = (int32_t*)calloc((size_t)W * H, sizeof(int32_t));
  int* br = (int*)malloc(sizeof(int) * (size_t)W * H);
  int* bk = (int*)calloc((size_t)W * H, sizeof(int));
  if (!m || !br || !bk) { free(m); free(br); free(bk); return INT32_MIN; }
  for (size_t q = 0; q < (size_t)W * H; q++) br[q] = -1;
  for (int jj = 0; jj < W; ++jj)
    for (int ii = jj; ii < H; ++ii) {
      const int i = H - 1 - ii, j = i + jj;
      int32_t c0 = 0; int rule = -1, pk = 0;
      if (i + 1 == j) { c0 = 0; rule = 0; }
      for (int k = i + 1; k < j; ++k) {
        /* JVM Int arithmetic: two's-complement wrap */
        uint32_t prod = (uint32_t)rows[i] * (uint32_t)rows[k] * (uint32_t)cols[j - 1];
        int32_t c = (int32_t)((uint32_t)m[IX(i, k)] + (uint32_t)m[IX(k, j)] + prod);
        if (c < c0 || rule == -1) { c0 = c; rule = 1; pk = k; }
      }
      m[IX(i, j)] = c0; br[IX(i, j)] = rule; bk[IX(i, j)] = pk;
    }
  const int32_t res = m[IX(0, W - 1)];
  if (trace && trace_len) {
    int rd = 0, wr = 0;
#define PUSH(a, b) do { if (wr < trace_cap) { trace[wr].i = (a); trace[wr].j = (b); trace[wr].rule = 0; trace[wr].pos[0] = trace[wr].pos[1] = 0; } wr++; } while (0)
    PUSH(0, W - 1);
    for (; rd < wr && rd < trace_cap; ++rd) {
      orc_trace_t* t = &trace[rd];
      t->rule = br[IX(t->i, t->j)];
      t->pos[0] = t->rule == 1 ? bk[IX(t->i, t->j)] : 0;
      if (t->rule == 1) { PUSH(t->i, t->pos[0]); PUSH(t->pos[0], t->j); }
    }
    *trace_len = wr;
#undef PUSH
  }
  free(m); free(br); free(bk);
  return res;
}
