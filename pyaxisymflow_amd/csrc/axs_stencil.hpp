// Launchers of the stencil kernels for the device-resident steps (the
// drop-in C entry points in axs_stencil.hip use the same kernels).
#pragma once

#include "axs_common.hpp"

namespace axs {

// smooth_Heaviside.py:5-14 at one point: 1 above w, 0 below -w, the sine
// blend between.
template <typename T>
__device__ __forceinline__ T smooth_heaviside(T p, T w) {
  constexpr double kPiH = 3.14159265358979323846;
  T h = (p >= w) ? T(1) : T(0);
  if (fabs(p) < w) h = h + T(0.5) * (T(1) + p / w + sin(T(kPiH) * p / w) / T(kPiH));
  return h;
}

}  // namespace axs

namespace axs {

// diffusion_RK2.py:4-44: w_out = RK2 step of w_in (border copied), tmp is
// the stage-1 field.
template <typename T, class RA>
hipError_t launch_rk2(const T* w_in, T* tmp, T* w_out, RA R, int nr, int nz, Scalar<T> half_nu_dt,
                      Scalar<T> nu_dt, T dx, hipStream_t s);
// The two stages of launch_rk2 as separate launches (the periodic form
// exchanges the ghost columns of tmp between them, diffusion_RK2.py:47-90).
template <typename T, class RA>
hipError_t launch_rk2_stage1(const T* w_in, T* tmp, RA R, int nr, int nz, Scalar<T> half_nu_dt,
                             T dx, hipStream_t s);
template <typename T, class RA>
hipError_t launch_rk2_stage2(const T* w_in, const T* tmp, T* w_out, RA R, int nr, int nz,
                             Scalar<T> nu_dt, T dx, hipStream_t s);
// compute_velocity_from_psi.py:4-17 with r per row.
template <typename T>
hipError_t launch_velocity_from_psi(T* u_z, T* u_r, const T* psi, const double* r, int nr, int nz,
                                    T dx, hipStream_t s);
// smooth_Heaviside.py:5-14 over n values.
template <typename T>
hipError_t launch_heaviside(T* H, const T* phi, T w, int64_t n, hipStream_t s);
// solid_sigma.py:4-29 (gradient scratch persists between calls); with chi,
// each stress is then multiplied by chi (the blend of
// soft_sphere_in_taylor_green_vortex.py:197-199) as it is stored.
template <typename T>
hipError_t launch_solid_sigma(T* s11, T* s12, T* s22, T G, T dx, const T* e1, const T* e2, T* e1z,
                              T* e1r, T* e2z, T* e2r, int nr, int nz, hipStream_t s,
                              const T* chi = nullptr);
// div_tau.py:4-34: tau (scratch persists), then omega's interior += dt curl tau.
template <typename T>
hipError_t launch_div_tau(T* w, T* tz, T* tr, const T* t11, const T* t12, const T* t22,
                          const double* r, int nr, int nz, Scalar<T> dt, T dx, hipStream_t s);
// Its two stages apart (the periodic form, div_tau.py:37-75, exchanges the
// ghost columns of tau between them).
template <typename T>
hipError_t launch_div_tau_stage1(T* tz, T* tr, const T* t11, const T* t12, const T* t22,
                                 const double* r, int nr, int nz, T dx, hipStream_t s);
template <typename T>
hipError_t launch_div_tau_stage2(T* w, const T* tz, const T* tr, int nr, int nz, Scalar<T> dt,
                                 T dx, hipStream_t s);

}  // namespace axs
