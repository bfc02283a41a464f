/*
 * ref_harness.c -- driver for the reference's own generated CPU code
 * (bench2/zu.c + zu_kern.h, bench2/nu.c + nu_kern.h), compiled from the
 * reference tree by oracle/Makefile.  TEST INFRASTRUCTURE ONLY.
 *
 * Same role as the reference harness bench2/wrapper.c:22-70, but it feeds
 * given sequences instead of random() ones.  Reads one sequence per stdin
 * line (letters ACGU, any case; its length must equal SIZE) and prints one
 * JSON object per line: {"n":N,"score":S,"trace":[[i,j,rule,p0,p1],...]}.
 * With argv[1] == "time" it prints {"n":N,"ms":T} for solve+backtrack.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>
#include <sys/time.h>
#ifdef NU
#include "nu.h"
#define TL 1
#else
#include "zu.h"
#define TL 2
#endif

int main(int argc, char** argv) {
  int timing = argc > 1 && strcmp(argv[1], "time") == 0;
#ifndef NU
  FILE* chk = fopen(LIBRNA "vienna/rna_turner2004.par", "r");
  if (!chk) { fprintf(stderr, "parameter file missing under %s\n", LIBRNA); return 2; }
  fclose(chk);
#endif
  char* line = (char*)malloc(SIZE + 64);
  char* seq = (char*)malloc(SIZE + 1);
  while (fgets(line, SIZE + 64, stdin)) {
    int n = 0;
    for (char* p = line; *p && *p != '\n'; ++p) {
      char c = (char)tolower((unsigned char)*p);
      if (n >= SIZE) { n = SIZE + 1; break; }
#ifdef NU
      seq[n++] = c;               /* nu_kern compares lower-case letters */
#else
      seq[n++] = c == 'a' ? 1 : c == 'c' ? 2 : c == 'g' ? 3 : c == 'u' ? 4 : 5;
#endif
    }
    if (n != SIZE) { fprintf(stderr, "sequence length %d != SIZE %d\n", n, SIZE); return 3; }
    struct timeval ts, te;
    my_init(seq, NULL);
    trace_t* trace = NULL;
    unsigned size = 0;
    gettimeofday(&ts, NULL);
    my_solve();
    int res = my_backtrack(&trace, &size);
    gettimeofday(&te, NULL);
    if (timing) {
      printf("{\"n\":%d,\"score\":%d,\"ms\":%.3f}\n", SIZE, res,
             (te.tv_sec - ts.tv_sec) * 1000.0 + (te.tv_usec - ts.tv_usec) / 1000.0);
    } else {
      printf("{\"n\":%d,\"score\":%d,\"trace\":[", SIZE, res);
      for (unsigned q = 0; q < size; ++q) {
        int r = trace[q].rule, tl = trace_len[r];
        printf("%s[%d,%d,%d,%d,%d]", q ? "," : "", trace[q].i, trace[q].j, r,
               tl > 0 ? trace[q].pos[0] : 0, TL > 1 && tl > 1 ? trace[q].pos[TL - 1] : 0);
      }
      printf("]}\n");
    }
    fflush(stdout);
    my_free();
    free(trace);
  }
  free(line); free(seq);
  return 0;
}
