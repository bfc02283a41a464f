/*
 * dpx_compat.h -- the reference's generated per-grammar C ABI on top of dpx.
 *
 * The reference's native libraries export, per grammar and baked-in size,
 *   void my_init(input_t* in1, input_t* in2); void my_solve();
 *   int my_backtrack(trace_t** trace, unsigned* size); void my_free();
 * (CodeGen.scala:291-358; bench2/zu.c:95-111, bench2/nu.c:61-73).  A harness
 * written against that ABI (e.g. bench2/wrapper.c) builds unchanged against
 * libdpx.so by including this header in one translation unit with
 *   #define DPX_COMPAT_ZUKER   (or DPX_COMPAT_NUSSINOV)
 *   #define SIZE <n>
 * input_t is char: Zuker takes base codes 1..4 (ADPParsers.scala:91 convert),
 * Nussinov takes the letters a,c,g,u.  Zuker loads the Turner 2004 parameters
 * from DPX_PARAMS (a path) at my_init, as the generated rna_init does.
 */
#ifndef DPX_COMPAT_H
#define DPX_COMPAT_H
#include <stdlib.h>
#include <string.h>
#include "dpx.h"

#if defined(DPX_COMPAT_ZUKER)
typedef struct { short i, j, rule; short pos[2]; } trace_t;
#define DPX_COMPAT_POS 2
#elif defined(DPX_COMPAT_NUSSINOV)
typedef struct { short i, j, rule; short pos[1]; } trace_t;
#define DPX_COMPAT_POS 1
#endif

#if defined(DPX_COMPAT_ZUKER) || defined(DPX_COMPAT_NUSSINOV)
#ifndef DPX_PARAMS
#define DPX_PARAMS "lamp-dp-mt_amd/params/rna_turner2004.par"
#endif
static dpx_ctx* dpx_compat_ctx = NULL;
static char dpx_compat_seq[SIZE + 1];
static int dpx_compat_score = 0;
static dpx_trace_t* dpx_compat_tr = NULL;
static int dpx_compat_tl = 0;

static void my_init(char* in1, char* in2) {
  (void)in2;
  static const char letters[] = "NACGU";
  for (int q = 0; q < SIZE; ++q) {
#if defined(DPX_COMPAT_ZUKER)
    dpx_compat_seq[q] = (in1[q] >= 1 && in1[q] <= 4) ? letters[(int)in1[q]] : 'N';
#else
    dpx_compat_seq[q] = in1[q];
#endif
  }
  if (!dpx_compat_ctx && dpx_create(&dpx_compat_ctx, 0) != DPX_OK) abort();
#if defined(DPX_COMPAT_ZUKER)
  if (dpx_set_rna_params(dpx_compat_ctx, DPX_PARAMS) != DPX_OK) abort();
#endif
}

static void my_solve(void) {
  const int cap = 4 * SIZE + 8;
  free(dpx_compat_tr);
  dpx_compat_tr = (dpx_trace_t*)malloc(sizeof(dpx_trace_t) * cap);
#if defined(DPX_COMPAT_ZUKER)
  int rc = dpx_zuker(dpx_compat_ctx, dpx_compat_seq, SIZE, &dpx_compat_score, dpx_compat_tr, cap, &dpx_compat_tl);
#else
  int rc = dpx_nussinov(dpx_compat_ctx, dpx_compat_seq, SIZE, 0, &dpx_compat_score, dpx_compat_tr, cap, &dpx_compat_tl);
#endif
  if (rc != DPX_OK) abort();
}

static int my_backtrack(trace_t** trace, unsigned* size) {
  if (trace && size) {
    *trace = (trace_t*)malloc(sizeof(trace_t) * (size_t)(dpx_compat_tl > 0 ? dpx_compat_tl : 1));
    for (int q = 0; q < dpx_compat_tl; ++q) {
      (*trace)[q].i = (short)dpx_compat_tr[q].i;
      (*trace)[q].j = (short)dpx_compat_tr[q].j;
      (*trace)[q].rule = (short)dpx_compat_tr[q].rule;
      for (int p = 0; p < DPX_COMPAT_POS; ++p) (*trace)[q].pos[p] = (short)dpx_compat_tr[q].pos[p];
    }
    *size = (unsigned)dpx_compat_tl;
  }
  return dpx_compat_score;
}

static void my_free(void) {
  free(dpx_compat_tr);
  dpx_compat_tr = NULL;
}
#endif
#endif
