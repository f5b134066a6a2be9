// Velocity potential plan (kernels/FastDiagonalisationPotentialSolver.py:6-145)
// shared by the drop-in entry points and the torus stepper.
#pragma once

#include "axs_common.hpp"
#include "axs_stokes.hpp"

namespace axs {

struct PotentialPlan {
  StokesPlan pl;  // DCT tables, the potential's column-pass tables, spectral buffer
  // mode 0 in the gauge phi_hat_0[nr - 1] = 0: bands with the last row
  // replaced, pivots at segment starts (L = 16) and the column buffer
  double *a0 = nullptr, *b0 = nullptr, *c0 = nullptr, *lam0 = nullptr, *cseg0 = nullptr;
  double* col0 = nullptr;
  hipError_t init(int nr, int nz, double dx);
  // phi = the solution of A_r phi + phi P_z^T = rhs in that gauge
  hipError_t solve(double* phi, const double* rhs, hipStream_t s);
  ~PotentialPlan();
};

// compute_velocity_from_phi.py:4-17 (u = grad phi, one-sided at the edges).
hipError_t launch_velocity_from_phi(double* u_z, double* u_r, const double* phi, int nr, int nz,
                                    double dx, hipStream_t s);

}  // namespace axs
