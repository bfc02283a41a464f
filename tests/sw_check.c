/*
 * sw_check.c -- independent checks of the local-alignment oracle (test
 * infrastructure; compiled by tests/test_oracle_pins.py with gcc).
 *
 * Reads lines "<mode> <a> <b> <match> <mismatch> <gap_open> <gap_ext>" ('-'
 * stands for an empty sequence) and prints "<best> <end>" per line:
 *   best = the best local alignment score (0 = the empty alignment),
 *   end  = the best score of a local alignment whose last column consumes
 *          a[n1-1] or b[n2-1] and nothing after them, i.e. that ends at
 *          (n1, n2); 0 when no alignment ending there is positive.
 * A gap run of length L costs gap_open + (L-1) * gap_ext.
 *
 * mode "brute": every alignment of every pair of substrings, enumerated
 *   column by column (no tables, no merging of partial alignments).
 * mode "gotoh": the textbook three-matrix recurrence (Gotoh 1982, in the
 *   form of Durbin et al. 1998, section 2.4) with the local zero only in H:
 *     E(i,j) = max(E(i,j-1) + ext, H(i,j-1) + open)
 *     F(i,j) = max(F(i-1,j) + ext, H(i-1,j) + open)
 *     H(i,j) = max(0, H(i-1,j-1) + s(a_i,b_j), E(i,j), F(i,j)),
 *   H(i,0) = H(0,j) = 0, E and F start at minus infinity; best = max H,
 *   end = H(n1,n2).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NEG (-(1 << 29))

static int sc_match, sc_mis, sc_open, sc_ext;
static const char *X, *Y;
static int LX, LY;

/* best score of a global alignment of X[i..LX) and Y[j..LY), the previous
 * column being prev (0 pair / start, 1 gap in Y (consumes X), 2 gap in X) */
static int rec(int i, int j, int prev) {
  if (i == LX && j == LY) return 0;
  int out = NEG;
  if (i < LX && j < LY) {
    int v = (X[i] == Y[j] ? sc_match : sc_mis) + rec(i + 1, j + 1, 0);
    if (v > out) out = v;
  }
  if (i < LX) {
    int v = (prev == 1 ? sc_ext : sc_open) + rec(i + 1, j, 1);
    if (v > out) out = v;
  }
  if (j < LY) {
    int v = (prev == 2 ? sc_ext : sc_open) + rec(i, j + 1, 2);
    if (v > out) out = v;
  }
  return out;
}

static void brute(const char* a, const char* b, int* best, int* end) {
  int n1 = (int)strlen(a), n2 = (int)strlen(b);
  *best = 0;
  *end = 0;
  for (int p = 0; p <= n1; ++p)
    for (int q = p; q <= n1; ++q)
      for (int r = 0; r <= n2; ++r)
        for (int s = r; s <= n2; ++s) {
          if (q == p && s == r) continue;
          X = a + p; LX = q - p; Y = b + r; LY = s - r;
          int v = rec(0, 0, 0);
          if (v > *best) *best = v;
          if (q == n1 && s == n2 && v > *end) *end = v;
        }
}

static void gotoh(const char* a, const char* b, int* best, int* end) {
  int n1 = (int)strlen(a), n2 = (int)strlen(b);
  int* H = calloc((size_t)(n2 + 1), sizeof(int));
  int* F = malloc(sizeof(int) * (size_t)(n2 + 1));
  for (int j = 0; j <= n2; ++j) F[j] = NEG;
  *best = 0;
  for (int i = 1; i <= n1; ++i) {
    int diag = H[0], e = NEG;  /* H(i-1,j-1), E(i,j-1) */
    H[0] = 0;
    for (int j = 1; j <= n2; ++j) {
      int up = H[j];  /* H(i-1,j) */
      int ev = e + sc_ext, eo = H[j - 1] + sc_open;
      e = ev > eo ? ev : eo;
      int fv = F[j] + sc_ext, fo = up + sc_open;
      F[j] = fv > fo ? fv : fo;
      int h = diag + (a[i - 1] == b[j - 1] ? sc_match : sc_mis);
      if (e > h) h = e;
      if (F[j] > h) h = F[j];
      if (h < 0) h = 0;
      H[j] = h;
      diag = up;
      if (h > *best) *best = h;
    }
  }
  *end = H[n2];
  free(H);
  free(F);
}

int main(void) {
  static char mode[16], a[1 << 16], b[1 << 16];
  while (scanf("%15s %65535s %65535s %d %d %d %d", mode, a, b, &sc_match, &sc_mis, &sc_open, &sc_ext) == 7) {
    if (strcmp(a, "-") == 0) a[0] = 0;
    if (strcmp(b, "-") == 0) b[0] = 0;
    int best, end;
    if (strcmp(mode, "brute") == 0) brute(a, b, &best, &end);
    else gotoh(a, b, &best, &end);
    printf("%d %d\n", best, end);
    fflush(stdout);
  }
  return 0;
}
