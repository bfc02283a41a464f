/*
 * dpx_oracle.h -- CPU oracle for the DynaProg DP hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * own CPU algorithms (the bench2 generated C of Nussinov/Zuker, the librna
 * energy model, and the generated SW/Gotoh and matrix-chain cell bodies).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the parity reference or the timed CPU baseline.  The product
 * (lamp-dp-mt_amd/) never links or calls it.
 *
 * Every routine cites the reference file:line it restates.  Parity status of
 * each routine is listed in oracle/README.md; all of them are pinned by at
 * least one independent check (reference golden vectors, brute force,
 * closed forms) in tests/test_oracle_*.py.
 */
#ifndef DPX_ORACLE_H
#define DPX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One backtrace record; same fields and meaning as the reference trace_t
 * (bench2/zu.h:17, nu.h:10, CodeGen.scala:239).  pos[] holds the split
 * indices of the winning rule (unused entries are 0). */
typedef struct { int32_t i, j, rule, pos[2]; } orc_trace_t;

/* Scaled Turner parameters, field-for-field the semantics of paramT
 * (src/librna/vienna/vienna.h:24-67) at 37 C. */
typedef struct {
  int16_t stack[8][8];
  int16_t hairpin[31], bulge[31], internal_loop[31];
  int16_t mismatchExt[8][5][5], mismatchI[8][5][5], mismatch1nI[8][5][5];
  int16_t mismatch23I[8][5][5], mismatchH[8][5][5], mismatchM[8][5][5];
  int16_t dangle5[8][5], dangle3[8][5];
  int32_t ninio[2];
  double lxc;
  int16_t MLbase, MLintern[8], MLclosing, TerminalAU, DuplexInit;
  int16_t Tetraloop_E[200]; char Tetraloops[1401]; int16_t TetraloopsLen;
  int16_t Triloop_E[40];    char Triloops[241];    int16_t TriloopsLen;
  int16_t Hexaloop_E[40];   char Hexaloops[1801];  int16_t HexaloopsLen;
  int16_t int11[8][8][5][5];
  int16_t int21[8][8][5][5][5];
  int16_t int22[8][8][5][5][5][5];
} orc_params_t;

/* Load a ViennaRNA v2.0 .par file and scale it (vienna.c:171-227, 13-96).
 * Returns 0 on success, nonzero if the file cannot be read or a section used
 * by the Zuker grammar is missing. */
int orc_load_params(const char* path, orc_params_t* out);

/* ---- energy functions on an encoded sequence (codes 1..4 = A,C,G,U) ---- */
typedef struct { const orc_params_t* P; const uint8_t* s; int n; } orc_rna_t;
int orc_bp_index(int x, int y);
int orc_hl_energy(const orc_rna_t* r, int i, int j);
int orc_bl_energy(const orc_rna_t* r, int i, int k, int l, int j, int xr);
int orc_br_energy(const orc_rna_t* r, int i, int k, int l, int j, int xl);
int orc_il_energy(const orc_rna_t* r, int i, int k, int l, int j);
int orc_sr_energy(const orc_rna_t* r, int i, int j);
int orc_termau_energy(const orc_rna_t* r, int i, int j);
int orc_ext_mismatch_energy(const orc_rna_t* r, int i, int j);
int orc_ml_mismatch_energy(const orc_rna_t* r, int i, int j);

/* ---- grammars ----
 * All return the axiom value.  If trace != NULL it must hold trace_cap
 * records; *trace_len receives the number written (BFS order, as the
 * reference cpu_backtrack).  Returns INT32_MIN on bad arguments. */

/* Zuker MFE (bench2/zu_kern.h:92-184, zu.c:39-111).  seq codes 1..4. */
int32_t orc_zuker(const orc_params_t* P, const uint8_t* seq, int n,
                  orc_trace_t* trace, int trace_cap, int* trace_len,
                  int32_t* tables /* optional (n+1)^2 x 4: cl,ml,ml1,st */);

/* RNAfold grammar with the MFE algebra (v4/examples/RNAFold.scala:20-64):
 *   Qp(i,j) [pair (i,j-1)] = min{ hairpin hl(i,j-1) [5]; stack sr(i,j-1)+Qp(i+1,j-1) [6];
 *                                 iloop il(i,k,l-1,j-1)+Qp(k,l), 1<=k-i<=30, 1<=j-l<=30 [7];
 *                                 mloop QM(i+1,j-1) [8] }
 *   QM(i,j) [j-i>=4] = min_k Q(i,k)+Q(k,j) [4]
 *   Q(i,j) [j-i>=2]  = min{ QM(i,j) [0]; Q(i,j-1) [1]; Q(i+1,j) [2]; Qp(i,j) [3] }
 * Rule ids as the reference's analyze numbers them (tables sorted by name:
 * Q 0-3, QM 4, Qp 5-8; BaseParsers.scala:140).  Answer Q(0,n); RF_EMPTY when
 * no structure exists.  Candidates in the order of the generated C code
 * (alternatives in order, split loops ascending, the outer split of
 * seq ~ Qp ~ seq (l) outside the inner one (k)); first optimum wins.
 * Trace: BFS from (0,n) like the other grammars; pos = (k, l) for rule 7, k
 * for rule 4.  tables (optional): (n+1)^2 x 3 = Q, QM, Qp. */
#define RF_EMPTY 999999
int32_t orc_rnafold(const orc_params_t* P, const uint8_t* seq, int n,
                    orc_trace_t* trace, int trace_cap, int* trace_len, int32_t* tables);

/* Nussinov (bench2/nu_kern.h:23-45, nu.c:17-73).  seq codes 1..4.
 * min_loop = 0 is the reference grammar (pair needs i+2<=j); min_loop = m
 * additionally requires m unpaired bases inside a pair. */
int32_t orc_nussinov(const uint8_t* seq, int n, int min_loop,
                     orc_trace_t* trace, int trace_cap, int* trace_len,
                     int32_t* table /* optional (n+1)^2 */);

/* SW with affine gaps, two-track (bench/dp_swat/comp1.cu:50-132,
 * SWatAffine.scala:22-69) with parametric scores.  Returns M(n1,n2);
 * *best receives max_ij M(i,j), *bi,*bj its first position in row-major
 * order.  If trace != NULL, traces from (n1,n2) when from_best == 0, else
 * from (*bi,*bj). */
int32_t orc_swat(const uint8_t* a, int n1, const uint8_t* b, int n2,
                 int match, int mismatch, int gap_open, int gap_ext,
                 int32_t* best, int* bi, int* bj, int from_best,
                 orc_trace_t* trace, int trace_cap, int* trace_len);

/* Matrix chain, Int variant (Benchmarks.scala:18-27, MatrixMult2.scala:16-27):
 * M(i,j) = min_k M(i,k)+M(k,j) + r[i]*r[k]*c[j-1], int32 two's-complement
 * wrap (JVM Int semantics).  Returns M(0,n). */
int32_t orc_matchain(const int32_t* rows, const int32_t* cols, int n,
                     orc_trace_t* trace, int trace_cap, int* trace_len);

#ifdef __cplusplus
}
#endif
#endif
