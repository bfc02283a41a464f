/*
 * dpx_compat.h -- the reference's generated per-grammar C ABI over libdpx.so.
 *
 * The reference's native backend emits, per grammar and baked-in size, one
 * C/CUDA file exporting
 *     void my_init(input_t* in1, input_t* in2); void my_solve();
 *     <Answer> my_backtrack(trace_t** trace, unsigned* size); void my_free();
 * (src/main/scala/v4/CodeGen.scala:291-358), together with a header giving
 * input_t, trace_t and trace_len[] (bench2/zu.h, bench2/nu.h,
 * bench/dp_swat/comp1.h).  A harness (bench2/wrapper.c) is one more
 * translation unit calling those four functions.
 *
 * This header is both halves over dpx.h.  Select the grammar and size with
 * macros before including it:
 *     DPX_COMPAT_ZUKER      bench2/zu.h:  input_t char (base codes 1..4,
 *                           ADPParsers.scala:91), trace_t pos[2], rules 0-14
 *     DPX_COMPAT_NUSSINOV   bench2/nu.h:  input_t char (letters a,c,g,u),
 *                           trace_t pos[1], rules 0-4
 *     DPX_COMPAT_RNAFOLD    RNAFold.scala: input_t char (codes 1..4),
 *                           trace_t pos[2], rules 0-8
 *     DPX_COMPAT_MATCHAIN   Benchmarks.scala:18-27: input_t T2ii {_1 rows,
 *                           _2 cols}, trace_t pos[1], rules 0-1
 *     DPX_COMPAT_SWAT       bench/dp_swat/comp1.h: two tracks, input_t char
 *                           (letters), trace_t without pos, rules 0-9;
 *                           scores of SWatAlgebra (SWatAffine.scala:50-69:
 *                           +10/-3, open -3, extend -1) unless DPX_SW_MATCH,
 *                           DPX_SW_MISMATCH, DPX_SW_OPEN, DPX_SW_EXT are set;
 *                           comp1.h's names g_init/g_solve/g_backtrack/g_free
 *                           are exported as well
 *     SIZE                  the sequence / chain length (bench2's -DSIZE); for
 *                           SW the lengths are M_H-1 and M_W-1 as in
 *                           comp1.cu:24-25 (default SIZE+1 each)
 * With DPX_COMPAT_DECLS_ONLY only the declarations are emitted (for the
 * harness's own translation unit when the definitions live in another one,
 * as lamp-dp-mt_amd/compat/<grammar>.c).  The definitions have external
 * linkage, so a harness compiled with the reference's own header links
 * against them unchanged.
 *
 * Behaviour differences: errors abort() with dpx_last_error on stderr (the
 * reference exit()s, CodeGen.scala:261); my_free keeps the device context
 * for the next my_init instead of resetting the device.  Zuker and RNAfold
 * read the Turner 2004 file at DPX_PARAMS (default: the repository-relative
 * lamp-dp-mt_amd/params/rna_turner2004.par), as the generated rna_init does.
 */
#ifndef DPX_COMPAT_H
#define DPX_COMPAT_H
#include "dpx.h"

#if defined(DPX_COMPAT_ZUKER) || defined(DPX_COMPAT_RNAFOLD)
#define input_t char
typedef struct { short i, j, rule; short pos[2]; } trace_t;
#define DPX_COMPAT_POS 2
#elif defined(DPX_COMPAT_NUSSINOV)
#define input_t char
typedef struct { short i, j, rule; short pos[1]; } trace_t;
#define DPX_COMPAT_POS 1
#elif defined(DPX_COMPAT_MATCHAIN)
typedef struct __T2ii T2ii;
struct __T2ii { int _1; int _2; };
#define input_t T2ii
typedef struct { short i, j, rule; short pos[1]; } trace_t;
#define DPX_COMPAT_POS 1
#elif defined(DPX_COMPAT_SWAT)
#define input_t char
typedef struct { short i, j, rule; } trace_t;
#define DPX_COMPAT_POS 0
#ifndef M_H
#define M_H (SIZE + 1)
#endif
#ifndef M_W
#define M_W (SIZE + 1)
#endif
#else
#error "define one of DPX_COMPAT_ZUKER, _NUSSINOV, _RNAFOLD, _MATCHAIN, _SWAT before including dpx_compat.h"
#endif

/* split positions recorded per rule: the maximum of its tabulation, as the
 * reference's headers list them (bench2/zu.h:18, nu.h:11, comp1.h:12) */
#if defined(DPX_COMPAT_ZUKER)
static const unsigned trace_len[15] __attribute__((unused)) = {2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1};
#elif defined(DPX_COMPAT_RNAFOLD)
static const unsigned trace_len[9] __attribute__((unused)) = {0, 0, 0, 0, 1, 2, 2, 2, 2};
#elif defined(DPX_COMPAT_NUSSINOV)
static const unsigned trace_len[5] __attribute__((unused)) = {1, 1, 1, 1, 1};
#elif defined(DPX_COMPAT_MATCHAIN)
static const unsigned trace_len[2] __attribute__((unused)) = {1, 1};
#else
static const unsigned trace_len[10] __attribute__((unused)) = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#endif

#ifdef __cplusplus
extern "C" {
#endif
void my_init(input_t* in1, input_t* in2);
void my_free();
void my_solve();
int my_backtrack(trace_t** trace, unsigned* size);
#if defined(DPX_COMPAT_SWAT)
void g_init(input_t* in1, input_t* in2);
void g_free();
void g_solve();
int g_backtrack(trace_t** trace, unsigned* size);
#endif
#ifdef __cplusplus
}
#endif

#ifndef DPX_COMPAT_DECLS_ONLY
#include <stdio.h>
#include <stdlib.h>

#if defined(DPX_COMPAT_SWAT)
#ifndef DPX_SW_MATCH
#define DPX_SW_MATCH 10
#endif
#ifndef DPX_SW_MISMATCH
#define DPX_SW_MISMATCH (-3)
#endif
#ifndef DPX_SW_OPEN
#define DPX_SW_OPEN (-3)
#endif
#ifndef DPX_SW_EXT
#define DPX_SW_EXT (-1)
#endif
#define DPX_COMPAT_N1 (M_H - 1)
#define DPX_COMPAT_N2 (M_W - 1)
#else
#define DPX_COMPAT_N1 (SIZE)
#define DPX_COMPAT_N2 0
#endif

#ifndef DPX_PARAMS
#define DPX_PARAMS "lamp-dp-mt_amd/params/rna_turner2004.par"
#endif

static dpx_ctx* dpx_compat_ctx = NULL;
#if !defined(DPX_COMPAT_MATCHAIN)
static char dpx_compat_a[DPX_COMPAT_N1 + 1];
#endif
#if defined(DPX_COMPAT_SWAT)
static char dpx_compat_b[DPX_COMPAT_N2 + 1];
#endif
#if defined(DPX_COMPAT_MATCHAIN)
static int32_t dpx_compat_rows[DPX_COMPAT_N1 + 1], dpx_compat_cols[DPX_COMPAT_N1 + 1];
#endif
static int32_t dpx_compat_score = 0;
static dpx_trace_t* dpx_compat_tr = NULL;
static int dpx_compat_tl = 0;

static void dpx_compat_check(int rc, const char* what) {
  if (rc != DPX_OK) {
    fprintf(stderr, "dpx_compat: %s failed (%d): %s\n", what, rc,
            dpx_compat_ctx ? dpx_last_error(dpx_compat_ctx) : "no context");
    abort();
  }
}

static void dpx_compat_close(void) {
  if (dpx_compat_ctx) dpx_destroy(dpx_compat_ctx);
  dpx_compat_ctx = NULL;
}

void my_init(input_t* in1, input_t* in2) {
  if (!dpx_compat_ctx) {
    dpx_compat_check(dpx_create(&dpx_compat_ctx, 0), "dpx_create");
    atexit(dpx_compat_close);
#if defined(DPX_COMPAT_ZUKER) || defined(DPX_COMPAT_RNAFOLD)
    dpx_compat_check(dpx_set_rna_params(dpx_compat_ctx, DPX_PARAMS), "dpx_set_rna_params");
#endif
  }
#if defined(DPX_COMPAT_ZUKER) || defined(DPX_COMPAT_RNAFOLD)
  static const char letters[] = "NACGU";  /* RNASignature.convert codes back to letters */
  for (int q = 0; q < DPX_COMPAT_N1; ++q)
    dpx_compat_a[q] = (in1[q] >= 1 && in1[q] <= 4) ? letters[(int)in1[q]] : 'N';
  (void)in2;
#elif defined(DPX_COMPAT_NUSSINOV)
  for (int q = 0; q < DPX_COMPAT_N1; ++q) dpx_compat_a[q] = in1[q];
  (void)in2;
#elif defined(DPX_COMPAT_MATCHAIN)
  for (int q = 0; q < DPX_COMPAT_N1; ++q) {
    dpx_compat_rows[q] = in1[q]._1;
    dpx_compat_cols[q] = in1[q]._2;
  }
  (void)in2;
#else
  for (int q = 0; q < DPX_COMPAT_N1; ++q) dpx_compat_a[q] = in1[q];
  for (int q = 0; q < DPX_COMPAT_N2; ++q) dpx_compat_b[q] = in2[q];
#endif
}

void my_solve() {
  /* trace capacity: an SW trace visits at most two table entries per
   * alignment column, a triangular grammar's at most 4 n + 8 */
  const int cap = 4 * (DPX_COMPAT_N1 + DPX_COMPAT_N2) + 8;
  free(dpx_compat_tr);
  dpx_compat_tr = (dpx_trace_t*)malloc(sizeof(dpx_trace_t) * (size_t)cap);
  if (!dpx_compat_tr) dpx_compat_check(DPX_ENOMEM, "malloc");
#if defined(DPX_COMPAT_ZUKER)
  dpx_compat_check(dpx_zuker(dpx_compat_ctx, dpx_compat_a, DPX_COMPAT_N1, &dpx_compat_score, dpx_compat_tr, cap,
                             &dpx_compat_tl), "dpx_zuker");
#elif defined(DPX_COMPAT_RNAFOLD)
  dpx_compat_check(dpx_rnafold(dpx_compat_ctx, dpx_compat_a, DPX_COMPAT_N1, &dpx_compat_score, dpx_compat_tr, cap,
                               &dpx_compat_tl), "dpx_rnafold");
#elif defined(DPX_COMPAT_NUSSINOV)
  dpx_compat_check(dpx_nussinov(dpx_compat_ctx, dpx_compat_a, DPX_COMPAT_N1, 0, &dpx_compat_score, dpx_compat_tr,
                                cap, &dpx_compat_tl), "dpx_nussinov");
#elif defined(DPX_COMPAT_MATCHAIN)
  dpx_compat_check(dpx_matchain(dpx_compat_ctx, dpx_compat_rows, dpx_compat_cols, DPX_COMPAT_N1, &dpx_compat_score,
                                dpx_compat_tr, cap, &dpx_compat_tl), "dpx_matchain");
#else
  {
    dpx_swat_scores sc = {DPX_SW_MATCH, DPX_SW_MISMATCH, DPX_SW_OPEN, DPX_SW_EXT};
    int32_t best, bi, bj;
    /* the reference's axiom: trace from (M_H-1, M_W-1) (comp1.cu:170) */
    dpx_compat_check(dpx_swat(dpx_compat_ctx, dpx_compat_a, DPX_COMPAT_N1, dpx_compat_b, DPX_COMPAT_N2, &sc,
                              &dpx_compat_score, &best, &bi, &bj, 0, dpx_compat_tr, cap, &dpx_compat_tl), "dpx_swat");
  }
#endif
}

int my_backtrack(trace_t** trace, unsigned* size) {
  if (trace && size) {
    *trace = (trace_t*)malloc(sizeof(trace_t) * (size_t)(dpx_compat_tl > 0 ? dpx_compat_tl : 1));
    if (!*trace) dpx_compat_check(DPX_ENOMEM, "malloc");
    for (int q = 0; q < dpx_compat_tl; ++q) {
      (*trace)[q].i = (short)dpx_compat_tr[q].i;
      (*trace)[q].j = (short)dpx_compat_tr[q].j;
      (*trace)[q].rule = (short)dpx_compat_tr[q].rule;
#if DPX_COMPAT_POS > 0
      for (int p = 0; p < DPX_COMPAT_POS; ++p) (*trace)[q].pos[p] = (short)dpx_compat_tr[q].pos[p];
#endif
    }
    *size = (unsigned)dpx_compat_tl;
  }
  return dpx_compat_score;
}

void my_free() {
  free(dpx_compat_tr);
  dpx_compat_tr = NULL;
  dpx_compat_tl = 0;
}

#if defined(DPX_COMPAT_SWAT)
void g_init(input_t* in1, input_t* in2) { my_init(in1, in2); }
void g_free() { my_free(); }
void g_solve() { my_solve(); }
int g_backtrack(trace_t** trace, unsigned* size) { return my_backtrack(trace, size); }
#endif
#endif /* DPX_COMPAT_DECLS_ONLY */
#endif /* DPX_COMPAT_H */
