/* Standalone run of the least-squares extrapolation oracle (lsq.c) for a
 * sanitizer build (test infrastructure only; tests/test_oracle_sanitized.py
 * compiles it with -fsanitize=address,undefined and runs it).
 *
 * Workload: the doubled (2 nr, nz) grid of a soft sphere's extrapolation,
 * z serving as both axes' coordinates as in the soft step: cells inside a
 * disc of radius 0.15 about (0.5, 0.5) known, those inside radius 0.15 +
 * band to be filled; eta_x = 0.3 + 2 x, eta_y = -0.1 + 0.5 y on the known
 * cells.  A first-order least-squares fit reproduces a linear field, so
 * every filled cell must hold it to round-off.  Exit status 0 when it does
 * and the loop took at least one sweep. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

int oracle_extrapolate_first_order(int16_t* current_flag, const int16_t* target_flag,
                                   double* eta_x, double* eta_y, const double* grid_x,
                                   const double* grid_y, int n0, int n1);

int main(void) {
  const int nr = 48, nz = 2 * nr, n0 = 2 * nr, n1 = nz;
  const double dx = 1.0 / nz, band = 8 * dx;
  const size_t n = (size_t)n0 * n1;
  int16_t* cur = calloc(n, sizeof(int16_t));
  int16_t* tgt = calloc(n, sizeof(int16_t));
  double* ex = calloc(n, sizeof(double));
  double* ey = calloc(n, sizeof(double));
  double* g = malloc(sizeof(double) * (size_t)n1);
  if (!cur || !tgt || !ex || !ey || !g) return 2;
  for (int i = 0; i < n1; ++i) g[i] = (i + 0.5) * dx;
  for (int j = 0; j < n0; ++j)
    for (int i = 0; i < n1; ++i) {
      const double x = g[i], y = (j + 0.5) * dx;
      const double d = sqrt((x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5));
      const size_t c = (size_t)j * n1 + i;
      cur[c] = d < 0.15;
      tgt[c] = d < 0.15 + band;
      if (cur[c]) {
        ex[c] = 0.3 + 2.0 * g[i];
        ey[c] = -0.1 + 0.5 * y;
      }
    }
  const int sweeps = oracle_extrapolate_first_order(cur, tgt, ex, ey, g, g, n0, n1);
  double worst = 0.0;
  long filled = 0;
  for (int j = 0; j < n0; ++j)
    for (int i = 0; i < n1; ++i) {
      const size_t c = (size_t)j * n1 + i;
      if (!tgt[c]) continue;
      if (!cur[c]) return 3; /* a target cell left unfilled */
      const double wx = 0.3 + 2.0 * g[i], wy = -0.1 + 0.5 * (j + 0.5) * dx;
      worst = fmax(worst, fmax(fabs(ex[c] - wx), fabs(ey[c] - wy)));
      ++filled;
    }
  printf("sweeps %d filled %ld max error %.3e\n", sweeps, filled, worst);
  free(cur);
  free(tgt);
  free(ex);
  free(ey);
  free(g);
  return (sweeps >= 1 && filled > 0 && worst < 1e-12) ? 0 : 1;
}
