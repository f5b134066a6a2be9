/*
 * C restatement of the reference's least-squares eta extrapolation
 * (TEST ORACLE ONLY; never linked into the product).
 *
 * Follows pyaxisymflow/core/src/extrapolate_using_least_squares.hpp:142-446
 * instantiated as extrapolate_using_least_squares_till_first_order
 * (:450-467: GridSize 3, ExtrapolationOrder 1, Dimensions 2), with
 *   - start codes  lstsq/ExtrapolationConfig.hpp:17-19, 131-158
 *   - 3x3 masks    lstsq/ExtrapolationConfig.hpp:31-51 (sum mask skips the
 *                  centre; x mask = column offset + 1; y mask = row offset + 1)
 *   - Gauss        lstsq/least_squares.hpp:10-59 (partial pivoting, double
 *                  abs as in the pybind11 build, where <stdlib.h> brings
 *                  std::abs(double) into the global namespace)
 * The arithmetic is in the same order as the reference, so on the same
 * compiler flags (no FMA contraction) the results agree bit for bit.
 *
 * Deviation (documented): the reference reads outside the arrays when a
 * candidate cell or its fitting window touches the array edge (no bounds
 * checks, .hpp:205-219, 276-280).  Here every out-of-range cell reads as
 * flag 0 (unknown), which contributes nothing to any sum or fit.
 *
 * Build: oracle/Makefile -> oracle/_lsq.so
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

typedef struct {
  int idx_x, idx_y;
  short start_x, start_y;
} point_info;

/* ExtrapolationConfig.hpp:17-19 */
static float round_up_to_next_even_number(float f) {
  return (float)(ceil(0.5 * f) * 2.0);
}

/* ExtrapolationConfig.hpp:131-158 with GridSize 3 (prefactor 2). */
static short start_code(short positional_sum, short total_sum) {
  float temp = (positional_sum > (short)4)
                   ? round_up_to_next_even_number((float)positional_sum)
                   : (float)positional_sum;
  short code = (short)(2.0 * temp / total_sum);
  return (short)((short)(2 * code - (short)6) / (short)2);
}

static short flag_at(const int16_t* f, int n0, int n1, int j, int i) {
  if (j < 0 || j >= n0 || i < 0 || i >= n1) return 0;
  return f[(int64_t)j * n1 + i];
}

/* least_squares.hpp:10-59 for NCoefficients 3, NComponents 2. */
static void gauss_elimination_3x2(double A[3][5], double sol[2][9]) {
  const short n_columns = 5;
  for (short i = 0; i < 3; i++) {
    short max_row = i;
    double max_element = fabs(A[i][i]);
    for (short k = i + 1; k < 3; k++) {
      if (fabs(A[k][i]) > max_element) {
        max_element = fabs(A[k][i]);
        max_row = k;
      }
    }
    for (short k = i; k < n_columns; k++) {
      double tmp = A[max_row][k];
      A[max_row][k] = A[i][k];
      A[i][k] = tmp;
    }
    for (short k = i + 1; k < 3; k++) {
      double c = -A[k][i] / A[i][i];
      A[k][i] = 0.0;
      for (short j = (short)(i + 1); j < n_columns; j++) A[k][j] += c * A[i][j];
    }
  }
  for (short comp = 0; comp < 2; ++comp) {
    for (short i = 2; i >= 0; --i) {
      sol[comp][i] = A[i][3 + comp] / A[i][i];
      for (short k = (short)(i - 1); k >= 0; --k)
        A[k][3 + comp] -= A[k][i] * sol[comp][i];
    }
  }
}

/* Exported for the oracle's own known-answer test (least_squares.hpp KAT). */
void oracle_gauss_elimination_3x2(double* A15, double* sol18) {
  double A[3][5], sol[2][9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 5; ++c) A[r][c] = A15[r * 5 + c];
  gauss_elimination_3x2(A, sol);
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 9; ++c) sol18[r * 9 + c] = sol[r][c];
}

/* Exported for the start-code table test. */
short oracle_start_code(short positional_sum, short total_sum) {
  return start_code(positional_sum, total_sum);
}

/* extrapolate_using_least_squares.hpp:142-446.  Returns the number of sweeps
 * that extrapolated at least one cell, or -1 on allocation failure. */
int oracle_extrapolate_first_order(int16_t* current_flag, const int16_t* target_flag,
                                   double* eta_x, double* eta_y, const double* grid_x,
                                   const double* grid_y, int n0, int n1) {
  point_info* info = (point_info*)malloc(sizeof(point_info) * (size_t)n0 * n1 + 1);
  if (!info) return -1;
  int sweeps = 0;
  for (;;) {
    int64_t count = 0;
    /* # 1. select cells (:197-253) */
    for (int j = 0; j < n0; ++j) {
      for (int i = 0; i < n1; ++i) {
        int64_t c = (int64_t)j * n1 + i;
        if (!(current_flag[c] ^ target_flag[c])) continue;
        short local_sum = 0;
        for (int jl = -1; jl <= 1; ++jl)
          for (int il = -1; il <= 1; ++il) {
            short m = (jl == 0 && il == 0) ? 0 : 1;
            local_sum += flag_at(current_flag, n0, n1, j + jl, i + il) * m;
          }
        if (!local_sum) continue;
        short lx = 0, ly = 0;
        for (int jl = -1; jl <= 1; ++jl)
          for (int il = -1; il <= 1; ++il) {
            short f = flag_at(current_flag, n0, n1, j + jl, i + il);
            lx += f * (short)(il + 1);
            ly += f * (short)(jl + 1);
          }
        info[count].idx_x = i;
        info[count].idx_y = j;
        info[count].start_x = start_code(lx, local_sum);
        info[count].start_y = start_code(ly, local_sum);
        ++count;
      }
    }
    if (!count) break;
    ++sweeps;
    /* # 3. fit and evaluate (:274-438) */
    for (int64_t p = 0; p < count; ++p) {
      const int s_y = info[p].idx_y + info[p].start_y;
      const int s_x = info[p].idx_x + info[p].start_x;
      double LHS[3][9], rhs[2][9], mask[9];
      int n = 0;
      for (int jl = s_y; jl < s_y + 3; ++jl) {
        for (int il = s_x; il < s_x + 3; ++il, ++n) {
          int in = (jl >= 0 && jl < n0 && il >= 0 && il < n1);
          rhs[0][n] = in ? grid_x[il] : 0.0;
          rhs[1][n] = in ? grid_y[jl] : 0.0;
          mask[n] = in ? (double)current_flag[(int64_t)jl * n1 + il] : 0.0;
        }
      }
      /* binomial_fill_data_parallel, order 1 (:67-85) */
      for (int q = 0; q < 9; ++q) LHS[0][q] = mask[q];
      for (int q = 0; q < 9; ++q) LHS[1][q] = LHS[0][q] * rhs[0][q];
      for (int q = 0; q < 9; ++q) LHS[2][q] = LHS[0][q] * rhs[1][q];
      n = 0;
      for (int jl = s_y; jl < s_y + 3; ++jl) {
        for (int il = s_x; il < s_x + 3; ++il, ++n) {
          int in = (jl >= 0 && jl < n0 && il >= 0 && il < n1);
          int64_t g = (int64_t)jl * n1 + il;
          rhs[0][n] = mask[n] * (in ? eta_x[g] : 0.0);
          rhs[1][n] = mask[n] * (in ? eta_y[g] : 0.0);
        }
      }
      /* multiply_by_transpose / _vec (:95-134) */
      double B[3][5];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          double s = 0.0;
          for (int q = 0; q < 9; ++q) s += LHS[a][q] * LHS[b][q];
          B[a][b] = s;
        }
      for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 2; ++d) {
          double s = 0.0;
          for (int q = 0; q < 9; ++q) s += LHS[a][q] * rhs[d][q];
          B[a][3 + d] = s;
        }
      gauss_elimination_3x2(B, rhs);
      /* evaluate at the cell (:416-436) */
      const double lx = grid_x[info[p].idx_x];
      const double ly = grid_y[info[p].idx_y];
      double L[3];
      L[0] = 1.0;
      L[1] = L[0] * lx;
      L[2] = L[0] * ly;
      const int64_t gc = (int64_t)info[p].idx_y * n1 + info[p].idx_x;
      double e = 0.0;
      for (int q = 0; q < 3; ++q) e += rhs[0][q] * L[q];
      eta_x[gc] = e;
      e = 0.0;
      for (int q = 0; q < 3; ++q) e += rhs[1][q] * L[q];
      eta_y[gc] = e;
    }
    /* # 4. flags after the sweep (:441-444) */
    for (int64_t p = 0; p < count; ++p)
      current_flag[(int64_t)info[p].idx_y * n1 + info[p].idx_x] = 1;
  }
  free(info);
  return sweeps;
}
