// Driver that compiles the reference's header-only least-squares extrapolation
// (pyaxisymflow/core/src/extrapolate_using_least_squares.hpp, read in place
// under /root/reference) into oracle/_ref/libref_lsq.so with a C ABI, so the
// oracle's C restatement can be checked against the reference itself.
// TEST INFRASTRUCTURE ONLY.
//
// <stdlib.h> comes first: in the reference's pybind11 build, Python.h includes
// it before the header, which puts std::abs(double) into the global namespace
// used by lstsq::gauss_elimination (least_squares.hpp:22-25).
#include <stdlib.h>

#include "extrapolate_using_least_squares.hpp"

extern "C" {

// extrapolate_using_least_squares.hpp:450-467, <short, double>, as bound by
// core/extrapolate_using_least_squares_bind.cpp:11-41.
void ref_extrapolate_first_order(short* current_flag, const short* target_flag,
                                 double* eta_x, double* eta_y,
                                 const double* grid_x, const double* grid_y,
                                 int n0, int n1) {
  extrapolate_using_least_squares_till_first_order<short, double>(
      current_flag, n0, n1, target_flag, n0, n1, eta_x, n0, n1, eta_y, n0, n1,
      grid_x, n1, grid_y, n0);
}

// lstsq/least_squares.hpp:10-59 with <3, 2, 9>, the instantiation the
// extrapolation uses (and the reference's own KAT, test_least_squares.cpp:43).
void ref_gauss_elimination_3x2(double* A15, double* sol18) {
  double A[3][5];
  double sol[2][9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 5; ++c) A[r][c] = A15[r * 5 + c];
  lstsq::gauss_elimination<3, 2, 9>(A, sol);
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 9; ++c) sol18[r * 9 + c] = sol[r][c];
}

// lstsq/ExtrapolationConfig.hpp:155-158 for the configuration above.
short ref_start_code(short positional_sum, short total_sum) {
  using EC = lstsq::extrapolation_configuration_t<
      lstsq::GridSize<3>, lstsq::ExtrapolationOrder<1>, lstsq::Dimensions<2>>;
  return EC::start_code<short>(positional_sum, total_sum);
}
}
