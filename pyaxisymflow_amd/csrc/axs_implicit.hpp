// Implicit Euler diffusion plan (kernels/implicit_diffusion_solver.py:6-143)
// shared by the drop-in stepper and the rigid stepper's implicit mode.
#pragma once

#include "axs_stokes.hpp"

namespace axs {

struct ImplicitPlan {
  StokesPlan tri;       // column-pass tables (diffusion bands) and the spectral buffer
  AnyDft dst;           // DFT of nz + 1 points for the sine transform rows
  Cplx<double>* em = nullptr;  // e^{-i pi k / (nz + 1)}, k <= nz
  double nu_dt = 0.0;
  // spectral buffer of batch (nr, nz) fields
  hipError_t init(int nr, int nz, double dx, double nu_dt, int batch = 1);
  // (I - nu dt L)^{-1} in place on batch (nr, nz) fields lying back to back
  hipError_t launch(double* omega, int batch, hipStream_t s);
  ~ImplicitPlan();
};

}  // namespace axs
