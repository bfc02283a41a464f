/*
 * harness.c -- a C harness written against the reference's generated ABI
 * (my_init / my_solve / my_backtrack / my_free, CodeGen.scala:291-358), in
 * the manner of bench2/wrapper.c but reading its inputs from stdin.  It
 * includes only the grammar's reference-shaped header (include/compat/
 * zu.h, nu.h, rf.h, mc.h or sw.h, chosen with -DZU, -DNU, -DRF, -DMC or
 * -DSW) and is linked with lamp-dp-mt_amd/compat/<grammar>.c and libdpx.so
 * by tests/test_gpu_compat.py.
 *
 * Input, one instance per line: letters ACGU (ZU, RF, NU; length SIZE), two
 * words a b (SW; lengths M_H-1, M_W-1), or 2*SIZE integers r0 c0 r1 c1 ...
 * (MC).  Output per instance: {"score":S,"trace":[[i,j,rule,p0,p1],...]}
 * with the positions trace_len[rule] lists (0 elsewhere).
 */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#if defined(ZU)
#include "zu.h"
#elif defined(NU)
#include "nu.h"
#elif defined(RF)
#include "rf.h"
#elif defined(MC)
#include "mc.h"
#else
#include "sw.h"
#endif

static void emit(int score, trace_t* tr, unsigned size) {
  printf("{\"score\":%d,\"trace\":[", score);
  for (unsigned q = 0; q < size; ++q) {
    int p0 = 0, p1 = 0;
#if !defined(SW)
    unsigned tl = trace_len[tr[q].rule];
    if (tl > 0) p0 = tr[q].pos[0];
#if defined(ZU) || defined(RF)
    if (tl > 1) p1 = tr[q].pos[1];
#endif
#endif
    printf("%s[%d,%d,%d,%d,%d]", q ? "," : "", tr[q].i, tr[q].j, tr[q].rule, p0, p1);
  }
  printf("]}\n");
  fflush(stdout);
}

int main(void) {
  static char line[1 << 20];
  while (fgets(line, sizeof line, stdin)) {
#if defined(SW)
    static char a[1 << 19], b[1 << 19];
    a[0] = b[0] = 0;
    if (sscanf(line, "%524287s %524287s", a, b) < 1) continue;
    if (strcmp(a, "-") == 0) a[0] = 0;
    if (strcmp(b, "-") == 0) b[0] = 0;
    if ((int)strlen(a) != M_H - 1 || (int)strlen(b) != M_W - 1) {
      fprintf(stderr, "lengths %zu x %zu != %d x %d\n", strlen(a), strlen(b), M_H - 1, M_W - 1);
      return 3;
    }
    my_init(a, b);
#elif defined(MC)
    static input_t in[SIZE + 1];
    char* p = line;
    for (int q = 0; q < SIZE; ++q) {
      in[q]._1 = (int)strtol(p, &p, 10);
      in[q]._2 = (int)strtol(p, &p, 10);
    }
    my_init(in, NULL);
#else
    static input_t in[SIZE + 1];
    int n = 0;
    for (char* p = line; *p && *p != '\n' && n <= SIZE; ++p) {
      char c = (char)tolower((unsigned char)*p);
#if defined(NU)
      in[n++] = c;  /* the reference's Nussinov compares letters (nu_kern.h) */
#else
      in[n++] = c == 'a' ? 1 : c == 'c' ? 2 : c == 'g' ? 3 : 4;  /* RNASignature.convert */
#endif
    }
    if (n != SIZE) {
      fprintf(stderr, "length %d != SIZE %d\n", n, SIZE);
      return 3;
    }
    my_init(in, NULL);
#endif
    my_solve();
    trace_t* tr = NULL;
    unsigned size = 0;
    int score = my_backtrack(&tr, &size);
    emit(score, tr, size);
    free(tr);
    my_free();
  }
  return 0;
}
