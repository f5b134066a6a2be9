// Pointwise and small-stencil kernels of the timestep (drop-in forms).
// Each kernel restates one reference function; the expression order follows
// the reference so FP64 results agree to round-off.
#include "axs_fused.hpp"
#include "axs_softbatch.hpp"
#include "axs_stencil.hpp"

#include <mutex>

namespace axs {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- A1
// kill_boundary_vorticity_sine.py:4-14: one thread per row; the left block
// is written first, then the right block reads column nz-width (as numpy does).
template <typename T>
__global__ void kill_z_kernel(T* __restrict__ w, const T* __restrict__ Z, int nr, int nz,
                              int width, T dx) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  T* row = w + (int64_t)i * nz;
  const T* zr = Z + (int64_t)i * nz;
  const T left = row[width - 1];
  for (int j = 0; j < width; ++j)
    row[j] = sin(T(kPi) * (zr[j] - T(0.5) * dx) / T(2) / T(width - 1) / dx) * left;
  const T right = row[nz - width];
  for (int j = nz - width; j < nz; ++j)
    row[j] = sin(T(kPi) * (T(1) - zr[j] - T(0.5) * dx) / T(2) / T(width - 1) / dx) * right;
}

// kill_boundary_vorticity_sine.py:17-27: one thread per column.
template <typename T>
__global__ void kill_r_kernel(T* __restrict__ w, const T* __restrict__ R, int nr, int nz,
                              int width, T dx) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nz) return;
  const T src = w[(int64_t)(nr - width) * nz + j];
  for (int i = nr - width; i < nr; ++i) {
    const int64_t c = (int64_t)i * nz + j;
    w[c] = sin(T(kPi) * (T(1) - R[c] - T(0.5) * dx) / T(2) / T(width - 1) / dx) * src;
  }
  w[j] = T(0);
}

// ---------------------------------------------------------------- A3
// compute_velocity_from_psi.py:4-17
template <typename T, class RA>
__device__ __forceinline__ void velocity_from_psi_at(T* __restrict__ u_z, T* __restrict__ u_r,
                                                     const T* __restrict__ psi, RA R, int nr,
                                                     int nz, T dx, int i, int j, T& uz_out,
                                                     T& ur_out) {
  const int64_t c = (int64_t)i * nz + j;
  const T two_dx = T(2) * dx;
  const T r = R.at(i, j);
  T uz;
  if (i == 0) {
    uz = (-psi[c + 2 * nz] + T(4) * psi[c + nz] - T(3) * psi[c]) / two_dx / r;
  } else if (i == nr - 1) {
    uz = (psi[c - 2 * nz] - T(4) * psi[c - nz] + T(3) * psi[c]) / two_dx / r;
  } else {
    uz = (psi[c + nz] - psi[c - nz]) / two_dx / r;
  }
  T ur;
  if (j == 0) {
    ur = -(-psi[c + 2] + T(4) * psi[c + 1] - T(3) * psi[c]) / two_dx / r;
  } else if (j == nz - 1) {
    ur = -(psi[c - 2] - T(4) * psi[c - 1] + T(3) * psi[c]) / two_dx / r;
  } else {
    ur = -(psi[c + 1] - psi[c - 1]) / two_dx / r;
  }
  u_z[c] = uz;
  u_r[c] = ur;
  uz_out = uz;
  ur_out = ur;
}

template <typename T, class RA>
__global__ void velocity_from_psi_kernel(T* __restrict__ u_z, T* __restrict__ u_r,
                                         const T* __restrict__ psi, RA R, int nr, int nz,
                                         T dx) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  int i = blockIdx.y;
  if (j >= nz) return;
  T uz, ur;
  velocity_from_psi_at<T, RA>(u_z, u_r, psi, R, nr, nz, dx, i, j, uz, ur);
}

// ---------------------------------------------------------------- A4
// compute_vorticity_from_velocity.py:4-13 (interior only)
template <typename T>
__global__ void vorticity_from_velocity_kernel(T* __restrict__ vort, const T* __restrict__ u_z,
                                               const T* __restrict__ u_r, int nr, int nz,
                                               T dx) {
  int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int i = blockIdx.y + 1;
  if (j >= nz - 1) return;
  const int64_t c = (int64_t)i * nz + j;
  const T two_dx = T(2) * dx;
  vort[c] = (u_r[c + 1] - u_r[c - 1]) / two_dx - (u_z[c + nz] - u_z[c - nz]) / two_dx;
}

// ---------------------------------------------------------------- A5
// brinkmann_penalize.py:4-16; lam*dt is formed first as in the reference.
template <typename T>
__global__ void brinkmann_kernel(T lam_dt, const T* __restrict__ chi, T U_z, T U_r,
                                 const T* __restrict__ gz, const T* __restrict__ gr,
                                 T* __restrict__ pz, T* __restrict__ pr, int64_t n) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const T ldc = lam_dt * chi[k];
  const T den = T(1) + ldc;
  pz[k] = (gz[k] + ldc * U_z) / den;
  pr[k] = (gr[k] + ldc * U_r) / den;
}

// ---------------------------------------------------------------- A6
// smooth_Heaviside.py:5-14
template <typename T>
__global__ void heaviside_kernel(T* __restrict__ H, const T* __restrict__ phi, T w,
                                 int64_t n) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  H[k] = smooth_heaviside(phi[k], w);
}

// ---------------------------------------------------------------- A8
// diffusion_RK2.py:4-44.  Stage 1 writes temp (= omega on the border);
// stage 2 updates omega's interior in place from temp.
// (f_{i+1} + f_{i-1} + f_{j+1} + f_{j-1} - 4 f) / dx^2 + (f_{i+1} - f_{i-1}) / (2dx) / r
//   - f * r^-2, in the reference's order (diffusion_RK2.py:15-25).
template <typename T>
__device__ __forceinline__ T axisym_laplacian(const T* __restrict__ f, int64_t c, int nz,
                                              T r, T dx2, T two_dx) {
  const T fc = f[c];
  const T up = f[c + nz], dn = f[c - nz];
  return (up + dn + f[c + 1] + f[c - 1] - T(4) * fc) / dx2 + (up - dn) / two_dx / r -
         fc * (T(1) / (r * r));
}

template <typename T, class RA>
__device__ __forceinline__ void rk2_stage1_at(const T* __restrict__ w, T* __restrict__ tmp, RA R,
                                              int nr, int nz, T half_nu_dt, T dx2, T two_dx,
                                              int i, int j) {
  const int64_t c = (int64_t)i * nz + j;
  if (i == 0 || i == nr - 1 || j == 0 || j == nz - 1) {
    tmp[c] = w[c];
    return;
  }
  tmp[c] = w[c] + half_nu_dt * axisym_laplacian<T>(w, c, nz, R.at(i, j), dx2, two_dx);
}

template <typename T, class RA>
__global__ void rk2_stage1_kernel(const T* __restrict__ w, T* __restrict__ tmp, RA R, int nr,
                                  int nz, Scalar<T> half_nu_dt_s, T dx2, T two_dx) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  int i = blockIdx.y;
  if (j >= nz) return;
  rk2_stage1_at<T, RA>(w, tmp, R, nr, nz, half_nu_dt_s.get(), dx2, two_dx, i, j);
}

template <typename T, class RA>
__global__ void rk2_stage2_kernel(T* __restrict__ w, const T* __restrict__ tmp, RA R, int nr,
                                  int nz, Scalar<T> nu_dt_s, T dx2, T two_dx) {
  int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int i = blockIdx.y + 1;
  if (j >= nz - 1) return;
  const int64_t c = (int64_t)i * nz + j;
  w[c] = w[c] + nu_dt_s.get() * axisym_laplacian<T>(tmp, c, nz, R.at(i, j), dx2, two_dx);
}

// Stage 2 writing a separate output: interior = src + nu dt L(tmp), border =
// src (so the field can return to its home buffer after an out-of-place step).
template <typename T, class RA>
__device__ __forceinline__ void rk2_stage2_out_at(const T* __restrict__ src,
                                                  const T* __restrict__ tmp, T* __restrict__ out,
                                                  RA R, int nr, int nz, T nu_dt, T dx2, T two_dx,
                                                  int i, int j) {
  const int64_t c = (int64_t)i * nz + j;
  if (i == 0 || i == nr - 1 || j == 0 || j == nz - 1) {
    out[c] = src[c];
    return;
  }
  out[c] = src[c] + nu_dt * axisym_laplacian<T>(tmp, c, nz, R.at(i, j), dx2, two_dx);
}

template <typename T, class RA>
__global__ void rk2_stage2_out_kernel(const T* __restrict__ src, const T* __restrict__ tmp,
                                      T* __restrict__ out, RA R, int nr, int nz,
                                      Scalar<T> nu_dt_s, T dx2, T two_dx) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  int i = blockIdx.y;
  if (j >= nz) return;
  rk2_stage2_out_at<T, RA>(src, tmp, out, R, nr, nz, nu_dt_s.get(), dx2, two_dx, i, j);
}

template <typename T, class RA>
hipError_t launch_rk2(const T* w_in, T* tmp, T* w_out, RA R, int nr, int nz, Scalar<T> half_nu_dt,
                      Scalar<T> nu_dt, T dx, hipStream_t s) {
  const T dx2 = dx * dx;
  rk2_stage1_kernel<T, RA><<<dim3(ceil_div(nz, 256), nr), 256, 0, s>>>(w_in, tmp, R, nr, nz,
                                                                       half_nu_dt, dx2, T(2) * dx);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  rk2_stage2_out_kernel<T, RA><<<dim3(ceil_div(nz, 256), nr), 256, 0, s>>>(
      w_in, tmp, w_out, R, nr, nz, nu_dt, dx2, T(2) * dx);
  return hipGetLastError();
}
template <typename T, class RA>
hipError_t launch_rk2_stage1(const T* w_in, T* tmp, RA R, int nr, int nz, Scalar<T> half_nu_dt,
                             T dx, hipStream_t s) {
  rk2_stage1_kernel<T, RA><<<dim3(ceil_div(nz, 256), nr), 256, 0, s>>>(w_in, tmp, R, nr, nz,
                                                                       half_nu_dt, dx * dx,
                                                                       T(2) * dx);
  return hipGetLastError();
}

template <typename T, class RA>
hipError_t launch_rk2_stage2(const T* w_in, const T* tmp, T* w_out, RA R, int nr, int nz,
                             Scalar<T> nu_dt, T dx, hipStream_t s) {
  rk2_stage2_out_kernel<T, RA><<<dim3(ceil_div(nz, 256), nr), 256, 0, s>>>(
      w_in, tmp, w_out, R, nr, nz, nu_dt, dx * dx, T(2) * dx);
  return hipGetLastError();
}
template hipError_t launch_rk2_stage1<double, RRow<double>>(const double*, double*, RRow<double>,
                                                            int, int, Scalar<double>, double,
                                                            hipStream_t);
template hipError_t launch_rk2_stage2<double, RRow<double>>(const double*, const double*, double*,
                                                            RRow<double>, int, int,
                                                            Scalar<double>, double, hipStream_t);
template hipError_t launch_rk2<double, RRow<double>>(const double*, double*, double*, RRow<double>,
                                                     int, int, Scalar<double>, Scalar<double>,
                                                     double, hipStream_t);
template hipError_t launch_rk2<float, RRow<float>>(const float*, float*, float*, RRow<float>, int,
                                                   int, Scalar<float>, Scalar<float>, float,
                                                   hipStream_t);

// ---------------------------------------------------------------- A11
// solid_sigma.py:4-29.  Gradient scratch entries the reference never writes
// keep their previous contents and still enter sigma.
// The reference map and the arithmetic in T; the gradient scratch, the
// mask and the stresses stored in O (O = T but in the FP32-bulk soft batch,
// whose reference map stays FP64).
template <typename T, typename O = T>
__device__ __forceinline__ void solid_sigma_at(O* __restrict__ s11, O* __restrict__ s12,
                                               O* __restrict__ s22, T G, T dx,
                                               const T* __restrict__ e1, const T* __restrict__ e2,
                                               O* __restrict__ e1z, O* __restrict__ e1r,
                                               O* __restrict__ e2z, O* __restrict__ e2r, int nr,
                                               int nz, const O* __restrict__ chi, int i, int j) {
  const int64_t c = (int64_t)i * nz + j;
  const T two_dx = T(2) * dx;
  T a1z = T(e1z[c]), a2z = T(e2z[c]), a1r = T(e1r[c]), a2r = T(e2r[c]);
  if (j >= 1 && j <= nz - 2) {
    a1z = (e1[c + 1] - e1[c - 1]) / two_dx;
    a2z = (e2[c + 1] - e2[c - 1]) / two_dx;
    e1z[c] = O(a1z);
    e2z[c] = O(a2z);
  }
  if (i == 0) {
    a1r = (-e1[c + 2 * nz] + T(4) * e1[c + nz] - T(3) * e1[c]) / two_dx;
    a2r = (-e2[c + 2 * nz] + T(4) * e2[c + nz] - T(3) * e2[c]) / two_dx;
    e1r[c] = O(a1r);
    e2r[c] = O(a2r);
  } else if (i <= nr - 2 && j >= 1 && j <= nz - 2) {
    a1r = (e1[c + nz] - e1[c - nz]) / two_dx;
    a2r = (e2[c + nz] - e2[c - nz]) / two_dx;
    e1r[c] = O(a1r);
    e2r[c] = O(a2r);
  }
  const T t12 = -G * (a1z * a1r + a2z * a2r);
  const T s = (T(0.5) * G) * (a1r * a1r + a2r * a2r - a1z * a1z - a2z * a2z);
  if (chi) {
    const T x = T(chi[c]);
    s12[c] = O(x * t12);
    s11[c] = O(x * s);
    s22[c] = O(x * -s);
  } else {
    s12[c] = O(t12);
    s11[c] = O(s);
    s22[c] = O(-s);
  }
}

template <typename T>
__global__ void solid_sigma_kernel(T* __restrict__ s11, T* __restrict__ s12,
                                   T* __restrict__ s22, T G, T dx,
                                   const T* __restrict__ e1, const T* __restrict__ e2,
                                   T* __restrict__ e1z, T* __restrict__ e1r,
                                   T* __restrict__ e2z, T* __restrict__ e2r, int nr, int nz,
                                   const T* __restrict__ chi) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  int i = blockIdx.y;
  if (j >= nz) return;
  solid_sigma_at<T>(s11, s12, s22, G, dx, e1, e2, e1z, e1r, e2z, e2r, nr, nz, chi, i, j);
}

// ---------------------------------------------------------------- A12
// div_tau.py:4-34: tau on the interior and on row 0 (one-sided in r) ...
template <typename T, class RA>
__device__ __forceinline__ void div_tau_stage1_at(T* __restrict__ tz, T* __restrict__ tr,
                                                  const T* __restrict__ t11,
                                                  const T* __restrict__ t12,
                                                  const T* __restrict__ t22, RA R, int nz,
                                                  T two_dx, int i, int j) {
  const int64_t c = (int64_t)i * nz + j;
  const T r = R.at(i, j);
  if (i == 0) {
    tz[c] = (t11[c + 1] - t11[c - 1] - t12[c + 2 * nz] + T(4) * t12[c + nz] - T(3) * t12[c]) /
                two_dx + t12[c] / r;
    tr[c] = (t12[c + 1] - t12[c - 1] - t22[c + 2 * nz] + T(4) * t22[c + nz] - T(3) * t22[c]) /
                two_dx + t22[c] / r;
  } else {
    tz[c] = (t11[c + 1] - t11[c - 1] + t12[c + nz] - t12[c - nz]) / two_dx + t12[c] / r;
    tr[c] = (t12[c + 1] - t12[c - 1] + t22[c + nz] - t22[c - nz]) / two_dx + t22[c] / r;
  }
}

template <typename T, class RA>
__global__ void div_tau_stage1_kernel(T* __restrict__ tz, T* __restrict__ tr,
                                      const T* __restrict__ t11, const T* __restrict__ t12,
                                      const T* __restrict__ t22, RA R, int nr, int nz,
                                      T two_dx) {
  int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int i = blockIdx.y;
  if (j >= nz - 1 || i >= nr - 1) return;
  div_tau_stage1_at<T, RA>(tz, tr, t11, t12, t22, R, nz, two_dx, i, j);
}

// ... then omega's interior gains dt * curl(tau).
template <typename T>
__device__ __forceinline__ void div_tau_stage2_at(T* __restrict__ w, const T* __restrict__ tz,
                                                  const T* __restrict__ tr, int nz, T dt,
                                                  T two_dx, int i, int j) {
  const int64_t c = (int64_t)i * nz + j;
  w[c] = w[c] + dt * (tr[c + 1] - tr[c - 1] - tz[c + nz] + tz[c - nz]) / two_dx;
}

template <typename T>
__global__ void div_tau_stage2_kernel(T* __restrict__ w, const T* __restrict__ tz,
                                      const T* __restrict__ tr, int nr, int nz, Scalar<T> dt_s,
                                      T two_dx) {
  int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int i = blockIdx.y + 1;
  if (j >= nz - 1) return;
  div_tau_stage2_at<T>(w, tz, tr, nz, dt_s.get(), two_dx, i, j);
}

// ---------------------------------------------------------------- soft batches
// One row segment of 256 cells of one case per workgroup (the batch's
// block table); columns beyond the case's own nz exit.

// u from psi, and the step's max(|u_z| + |u_r|) folded per workgroup into the
// case's spread atomicMax slots (the maximum does not depend on the order).
template <typename B>
__global__ void __launch_bounds__(256) soft_batch_velocity_kernel(
    const SoftCaseT<B>* __restrict__ cs, const uint32_t* __restrict__ blocks) {
  const SoftBlock b = soft_block(blocks);
  const SoftCaseT<B>& C = cs[b.c];
  const int i = b.i, j = b.j;
  double m = 0.0;
  if (j < C.nz) {
    B uz, ur;
    velocity_from_psi_at<B, RRow<B>>(C.u_z, C.u_r, C.psi, RRow<B>{C.r}, C.nr, C.nz, B(C.dx), i,
                                     j, uz, ur);
    m = double(fabs(uz) + fabs(ur));
  }
  // NaN-propagating maximum (a NaN velocity makes the case's dt NaN, as
  // numpy's amax does in the reference's dt rule)
  __shared__ double wm[4];
  for (int off = 32; off > 0; off >>= 1) m = max_nan(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = max_nan(max_nan(wm[0], wm[1]), max_nan(wm[2], wm[3]));
    const int slot = blockIdx.x % kMaxSlots;
    atomicMax(reinterpret_cast<unsigned long long*>(C.slots + slot * kSlotStride),
              static_cast<unsigned long long>(__double_as_longlong(m)));
  }
}

// Stresses from the FP64 reference map, stored in the batch's type.
template <typename B>
__global__ void soft_batch_sigma_kernel(const SoftCaseT<B>* __restrict__ cs,
                                        const uint32_t* __restrict__ blocks) {
  const SoftBlock b = soft_block(blocks);
  const SoftCaseT<B>& C = cs[b.c];
  const int i = b.i, j = b.j;
  if (j >= C.nz || !zone_row(C, i)) return;
  solid_sigma_at<double, B>(C.s11, C.s12, C.s22, C.G, C.dx, C.eta1, C.eta2, C.g1z, C.g1r, C.g2z,
                            C.g2r, C.nr, C.nz, C.chi, i, j);
}

template <typename B>
__global__ void soft_batch_div_tau1_kernel(const SoftCaseT<B>* __restrict__ cs,
                                           const uint32_t* __restrict__ blocks) {
  const SoftBlock b = soft_block(blocks);
  const SoftCaseT<B>& C = cs[b.c];
  const int i = b.i, j = b.j + 1;
  if (i >= C.nr - 1 || j >= C.nz - 1 || !zone_row(C, i)) return;
  div_tau_stage1_at<B, RRow<B>>(C.tz, C.tr, C.s11, C.s12, C.s22, RRow<B>{C.r}, C.nz,
                                B(2.0 * C.dx), i, j);
}

template <typename B>
__global__ void soft_batch_div_tau2_kernel(const SoftCaseT<B>* __restrict__ cs,
                                           const uint32_t* __restrict__ blocks) {
  const SoftBlock b = soft_block(blocks);
  const SoftCaseT<B>& C = cs[b.c];
  const int i = b.i, j = b.j + 1;
  if (i < 1 || i >= C.nr - 1 || j >= C.nz - 1 || !zone_row(C, i)) return;
  div_tau_stage2_at<B>(C.w0, C.tz, C.tr, C.nz, B(C.scalars[kSsDt]), B(2.0 * C.dx), i, j);
}

template <typename B>
__global__ void soft_batch_rk2_stage1_kernel(const SoftCaseT<B>* __restrict__ cs,
                                             const uint32_t* __restrict__ blocks) {
  const SoftBlock b = soft_block(blocks);
  const SoftCaseT<B>& C = cs[b.c];
  const int i = b.i, j = b.j;
  if (j >= C.nz) return;
  rk2_stage1_at<B, RRow<B>>(C.w1, C.tmp, RRow<B>{C.r}, C.nr, C.nz, B(C.scalars[kSsHalfNuDt]),
                            B(C.dx * C.dx), B(2.0 * C.dx), i, j);
}

template <typename B>
__global__ void soft_batch_rk2_stage2_kernel(const SoftCaseT<B>* __restrict__ cs,
                                             const uint32_t* __restrict__ blocks) {
  const SoftBlock b = soft_block(blocks);
  const SoftCaseT<B>& C = cs[b.c];
  const int i = b.i, j = b.j;
  if (j >= C.nz) return;
  rk2_stage2_out_at<B, RRow<B>>(C.w1, C.tmp, C.vorticity, RRow<B>{C.r}, C.nr, C.nz,
                                B(C.scalars[kSsNuDt]), B(C.dx * C.dx), B(2.0 * C.dx), i, j);
}

template <typename B>
hipError_t launch_soft_batch_velocity(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                      hipStream_t s) {
  soft_batch_velocity_kernel<B><<<d.ncell_blocks, 256, 0, s>>>(cases, d.cell_blocks);
  return hipGetLastError();
}

template <typename B>
hipError_t launch_soft_batch_stress(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                    hipStream_t s) {
  soft_batch_sigma_kernel<B><<<d.ncell_blocks, 256, 0, s>>>(cases, d.cell_blocks);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  soft_batch_div_tau1_kernel<B><<<d.ncell_blocks, 256, 0, s>>>(cases, d.cell_blocks);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  soft_batch_div_tau2_kernel<B><<<d.ncell_blocks, 256, 0, s>>>(cases, d.cell_blocks);
  return hipGetLastError();
}

template <typename B>
hipError_t launch_soft_batch_diffuse(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                     hipStream_t s) {
  soft_batch_rk2_stage1_kernel<B><<<d.ncell_blocks, 256, 0, s>>>(cases, d.cell_blocks);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  soft_batch_rk2_stage2_kernel<B><<<d.ncell_blocks, 256, 0, s>>>(cases, d.cell_blocks);
  return hipGetLastError();
}

#define AXS_SOFT_BATCH_INST(B)                                                              \
  template hipError_t launch_soft_batch_velocity<B>(const SoftCaseT<B>*, const SoftBatchDims&, \
                                                    hipStream_t);                           \
  template hipError_t launch_soft_batch_stress<B>(const SoftCaseT<B>*, const SoftBatchDims&,   \
                                                  hipStream_t);                             \
  template hipError_t launch_soft_batch_diffuse<B>(const SoftCaseT<B>*, const SoftBatchDims&,  \
                                                   hipStream_t);
AXS_SOFT_BATCH_INST(double)
AXS_SOFT_BATCH_INST(float)
#undef AXS_SOFT_BATCH_INST

// ---------------------------------------------------------------- launchers
// (device-resident steps pass dt-derived scalars from device memory)

template <typename T>
hipError_t launch_velocity_from_psi(T* u_z, T* u_r, const T* psi, const double* r, int nr, int nz,
                                    T dx, hipStream_t s) {
  velocity_from_psi_kernel<T, RRow<T>><<<dim3(ceil_div(nz, 256), nr), 256, 0, s>>>(
      u_z, u_r, psi, RRow<T>{r}, nr, nz, dx);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_heaviside(T* H, const T* phi, T w, int64_t n, hipStream_t s) {
  heaviside_kernel<T><<<ceil_div(n, 256), 256, 0, s>>>(H, phi, w, n);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_solid_sigma(T* s11, T* s12, T* s22, T G, T dx, const T* e1, const T* e2, T* e1z,
                              T* e1r, T* e2z, T* e2r, int nr, int nz, hipStream_t s,
                              const T* chi) {
  solid_sigma_kernel<T><<<dim3(ceil_div(nz, 256), nr), 256, 0, s>>>(
      s11, s12, s22, G, dx, e1, e2, e1z, e1r, e2z, e2r, nr, nz, chi);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_div_tau_stage1(T* tz, T* tr, const T* t11, const T* t12, const T* t22,
                                 const double* r, int nr, int nz, T dx, hipStream_t s) {
  div_tau_stage1_kernel<T, RRow<T>><<<dim3(ceil_div(nz - 2, 256), nr - 1), 256, 0, s>>>(
      tz, tr, t11, t12, t22, RRow<T>{r}, nr, nz, T(2) * dx);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_div_tau_stage2(T* w, const T* tz, const T* tr, int nr, int nz, Scalar<T> dt,
                                 T dx, hipStream_t s) {
  div_tau_stage2_kernel<T><<<dim3(ceil_div(nz - 2, 256), nr - 2), 256, 0, s>>>(w, tz, tr, nr, nz,
                                                                              dt, T(2) * dx);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_div_tau(T* w, T* tz, T* tr, const T* t11, const T* t12, const T* t22,
                          const double* r, int nr, int nz, Scalar<T> dt, T dx, hipStream_t s) {
  hipError_t e = launch_div_tau_stage1<T>(tz, tr, t11, t12, t22, r, nr, nz, dx, s);
  if (e != hipSuccess) return e;
  return launch_div_tau_stage2<T>(w, tz, tr, nr, nz, dt, dx, s);
}

template hipError_t launch_velocity_from_psi<double>(double*, double*, const double*,
                                                     const double*, int, int, double,
                                                     hipStream_t);
template hipError_t launch_heaviside<double>(double*, const double*, double, int64_t,
                                             hipStream_t);
template hipError_t launch_solid_sigma<double>(double*, double*, double*, double, double,
                                               const double*, const double*, double*, double*,
                                               double*, double*, int, int, hipStream_t,
                                               const double*);
template hipError_t launch_div_tau<double>(double*, double*, double*, const double*,
                                           const double*, const double*, const double*, int, int,
                                           Scalar<double>, double, hipStream_t);
template hipError_t launch_div_tau_stage1<double>(double*, double*, const double*, const double*,
                                                  const double*, const double*, int, int, double,
                                                  hipStream_t);
template hipError_t launch_div_tau_stage2<double>(double*, const double*, const double*, int, int,
                                                  Scalar<double>, double, hipStream_t);

}  // namespace axs

using namespace axs;

// periodic_boundary_ghost_comm.py:4-32: the ghost columns of a field
// periodic along z, field[:, :g] = field[:, -2g:-g] and
// field[:, -g:] = field[:, g:2g]; with eta_shift the reference map's
// version subtracts / adds the period, (x - Z_max) + 2 g dx on the left and
// (x + Z_max) - 2 g dx on the right (c = 2 g dx as the reference forms it).
// Sources and destinations are disjoint for nz >= 4 g.
__global__ void periodic_ghost_comm_kernel(double* __restrict__ f, int nz, int g, int eta_shift,
                                           double z_max, double c) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 2 * g) return;
  double* row = f + (int64_t)blockIdx.y * nz;
  if (q < g) {
    double v = row[nz - 2 * g + q];
    if (eta_shift) v = (v - z_max) + c;
    row[q] = v;
  } else {
    double v = row[q];  // column g + (q - g)
    if (eta_shift) v = (v + z_max) - c;
    row[nz - g + (q - g)] = v;
  }
}

static hipError_t launch_ghost_comm(double* f, int nr, int nz, int g, int eta_shift, double z_max,
                                    double dx, hipStream_t s) {
  const double c = (double)(2 * g) * dx;
  periodic_ghost_comm_kernel<<<dim3(ceil_div(2 * g, 64), nr), 64, 0, s>>>(f, nz, g, eta_shift,
                                                                          z_max, c);
  return hipGetLastError();
}

extern "C" {

const char* axs_last_error(void) { return g_last_error.c_str(); }
const char* axs_version(void) { return "axisym-hip 0.1 (gfx950)"; }

axs_status axs_kill_boundary_vorticity_sine_z_f64(double* vorticity, const double* Z, int nr,
                                                  int nz, int width, double dx,
                                                  void* stream) {
  AXS_CHECK_ARG(vorticity && Z, "null pointer");
  AXS_CHECK_ARG(width >= 2 && nz >= width && nr >= 1, "need width >= 2 and nz >= width");
  kill_z_kernel<double><<<ceil_div(nr, 64), 64, 0, as_stream(stream)>>>(vorticity, Z, nr, nz,
                                                                        width, dx);
  AXS_LAUNCH_CHECK("kill_z");
  return AXS_OK;
}

axs_status axs_kill_boundary_vorticity_sine_r_f64(double* vorticity, const double* R, int nr,
                                                  int nz, int width, double dx,
                                                  void* stream) {
  AXS_CHECK_ARG(vorticity && R, "null pointer");
  AXS_CHECK_ARG(width >= 2 && nr >= width && nz >= 1, "need width >= 2 and nr >= width");
  kill_r_kernel<double><<<ceil_div(nz, 256), 256, 0, as_stream(stream)>>>(vorticity, R, nr,
                                                                          nz, width, dx);
  AXS_LAUNCH_CHECK("kill_r");
  return AXS_OK;
}

axs_status axs_compute_velocity_from_psi_unb_f64(double* u_z, double* u_r, const double* psi,
                                                 const double* R, int nr, int nz, double dx,
                                                 void* stream) {
  AXS_CHECK_ARG(u_z && u_r && psi && R, "null pointer");
  AXS_CHECK_ARG(nr >= 3 && nz >= 3, "need nr >= 3 and nz >= 3");
  dim3 grid(ceil_div(nz, 256), nr);
  velocity_from_psi_kernel<double, RFull<double>><<<grid, 256, 0, as_stream(stream)>>>(
      u_z, u_r, psi, RFull<double>{R, nz}, nr, nz, dx);
  AXS_LAUNCH_CHECK("velocity_from_psi");
  return AXS_OK;
}

axs_status axs_compute_vorticity_from_velocity_unb_f64(double* vort, const double* u_z,
                                                       const double* u_r, int nr, int nz,
                                                       double dx, void* stream) {
  AXS_CHECK_ARG(vort && u_z && u_r, "null pointer");
  AXS_CHECK_ARG(nr >= 1 && nz >= 1, "bad shape");
  if (nr < 3 || nz < 3) return AXS_OK;  // no interior
  dim3 grid(ceil_div(nz - 2, 256), nr - 2);
  vorticity_from_velocity_kernel<double><<<grid, 256, 0, as_stream(stream)>>>(vort, u_z, u_r,
                                                                              nr, nz, dx);
  AXS_LAUNCH_CHECK("vorticity_from_velocity");
  return AXS_OK;
}

axs_status axs_brinkmann_penalize_f64(double lam, double dt, const double* char_func,
                                      double U_z, double U_r, const double* grid_u_z,
                                      const double* grid_u_r, double* penalized_u_z,
                                      double* penalized_u_r, int64_t n, void* stream) {
  AXS_CHECK_ARG(char_func && grid_u_z && grid_u_r && penalized_u_z && penalized_u_r,
                "null pointer");
  AXS_CHECK_ARG(n >= 0, "negative size");
  if (n == 0) return AXS_OK;
  brinkmann_kernel<double><<<ceil_div(n, 256), 256, 0, as_stream(stream)>>>(
      lam * dt, char_func, U_z, U_r, grid_u_z, grid_u_r, penalized_u_z, penalized_u_r, n);
  AXS_LAUNCH_CHECK("brinkmann");
  return AXS_OK;
}

axs_status axs_smooth_heaviside_f64(double* H, const double* phi, double blend_w, int64_t n,
                                    void* stream) {
  AXS_CHECK_ARG(H && phi, "null pointer");
  AXS_CHECK_ARG(n >= 0, "negative size");
  if (n == 0) return AXS_OK;
  heaviside_kernel<double><<<ceil_div(n, 256), 256, 0, as_stream(stream)>>>(H, phi, blend_w, n);
  AXS_LAUNCH_CHECK("heaviside");
  return AXS_OK;
}

axs_status axs_diffusion_rk2_unb_f64(double* vorticity, double* temp_vorticity,
                                     const double* R, int nr, int nz, double nu, double dt,
                                     double dx, void* stream) {
  AXS_CHECK_ARG(vorticity && temp_vorticity && R, "null pointer");
  AXS_CHECK_ARG(nr >= 1 && nz >= 1, "bad shape");
  AXS_CHECK_ARG(vorticity != temp_vorticity, "vorticity and temp_vorticity must differ");
  hipStream_t s = as_stream(stream);
  RFull<double> RA{R, nz};
  const double dx2 = dx * dx;
  rk2_stage1_kernel<double, RFull<double>><<<dim3(ceil_div(nz, 256), nr), 256, 0, s>>>(
      vorticity, temp_vorticity, RA, nr, nz, Scalar<double>{0.5 * nu * dt, nullptr}, dx2, 2.0 * dx);
  AXS_LAUNCH_CHECK("rk2_stage1");
  if (nr >= 3 && nz >= 3) {
    rk2_stage2_kernel<double, RFull<double>><<<dim3(ceil_div(nz - 2, 256), nr - 2), 256, 0, s>>>(
        vorticity, temp_vorticity, RA, nr, nz, Scalar<double>{nu * dt, nullptr}, dx2, 2.0 * dx);
    AXS_LAUNCH_CHECK("rk2_stage2");
  }
  return AXS_OK;
}

axs_status axs_solid_sigma_f64(double* sigma_s_11, double* sigma_s_12, double* sigma_s_22,
                               double G, double dx, const double* eta_1, const double* eta_2,
                               double* eta_1z, double* eta_1r, double* eta_2z, double* eta_2r,
                               int nr, int nz, void* stream) {
  AXS_CHECK_ARG(sigma_s_11 && sigma_s_12 && sigma_s_22 && eta_1 && eta_2 && eta_1z &&
                    eta_1r && eta_2z && eta_2r,
                "null pointer");
  AXS_CHECK_ARG(nr >= 3 && nz >= 3, "need nr >= 3 and nz >= 3");
  solid_sigma_kernel<double><<<dim3(ceil_div(nz, 256), nr), 256, 0, as_stream(stream)>>>(
      sigma_s_11, sigma_s_12, sigma_s_22, G, dx, eta_1, eta_2, eta_1z, eta_1r, eta_2z, eta_2r,
      nr, nz, nullptr);
  AXS_LAUNCH_CHECK("solid_sigma");
  return AXS_OK;
}

axs_status axs_update_vorticity_from_solid_stress_f64(double* vorticity, double* tau_z,
                                                      double* tau_r, const double* tau11,
                                                      const double* tau12,
                                                      const double* tau22, const double* R,
                                                      int nr, int nz, double dt, double dx,
                                                      void* stream) {
  AXS_CHECK_ARG(vorticity && tau_z && tau_r && tau11 && tau12 && tau22 && R, "null pointer");
  AXS_CHECK_ARG(nr >= 3 && nz >= 3, "need nr >= 3 and nz >= 3");
  hipStream_t s = as_stream(stream);
  div_tau_stage1_kernel<double, RFull<double>><<<dim3(ceil_div(nz - 2, 256), nr - 1), 256, 0, s>>>(
      tau_z, tau_r, tau11, tau12, tau22, RFull<double>{R, nz}, nr, nz, 2.0 * dx);
  AXS_LAUNCH_CHECK("div_tau_stage1");
  div_tau_stage2_kernel<double><<<dim3(ceil_div(nz - 2, 256), nr - 2), 256, 0, s>>>(
      vorticity, tau_z, tau_r, nr, nz, Scalar<double>{dt, nullptr}, 2.0 * dx);
  AXS_LAUNCH_CHECK("div_tau_stage2");
  return AXS_OK;
}

axs_status axs_periodic_ghost_comm_f64(double* field, int nr, int nz, int ghost, int eta_shift,
                                       double z_max, double dx, void* stream) {
  AXS_CHECK_ARG(field, "null pointer");
  AXS_CHECK_ARG(ghost > 0 && nr >= 1 && nz >= 4 * ghost, "need ghost > 0 and nz >= 4 ghost");
  AXS_HIP(launch_ghost_comm(field, nr, nz, ghost, eta_shift, z_max, dx, as_stream(stream)));
  return AXS_OK;
}

axs_status axs_diffusion_rk2_periodic_f64(double* vorticity, double* temp_vorticity,
                                          const double* R, int nr, int nz, double nu, double dt,
                                          double dx, int ghost, void* stream) {
  AXS_CHECK_ARG(vorticity && temp_vorticity && R, "null pointer");
  AXS_CHECK_ARG(nr >= 3 && ghost > 0 && nz >= 4 * ghost, "need nr >= 3, nz >= 4 ghost");
  AXS_CHECK_ARG(vorticity != temp_vorticity, "vorticity and temp_vorticity must differ");
  hipStream_t s = as_stream(stream);
  RFull<double> RA{R, nz};
  const double dx2 = dx * dx;
  AXS_HIP(launch_ghost_comm(vorticity, nr, nz, ghost, 0, 0.0, dx, s));
  rk2_stage1_kernel<double, RFull<double>><<<dim3(ceil_div(nz, 256), nr), 256, 0, s>>>(
      vorticity, temp_vorticity, RA, nr, nz, Scalar<double>{0.5 * nu * dt, nullptr}, dx2, 2.0 * dx);
  AXS_LAUNCH_CHECK("rk2_stage1");
  AXS_HIP(launch_ghost_comm(temp_vorticity, nr, nz, ghost, 0, 0.0, dx, s));
  rk2_stage2_kernel<double, RFull<double>><<<dim3(ceil_div(nz - 2, 256), nr - 2), 256, 0, s>>>(
      vorticity, temp_vorticity, RA, nr, nz, Scalar<double>{nu * dt, nullptr}, dx2, 2.0 * dx);
  AXS_LAUNCH_CHECK("rk2_stage2");
  return AXS_OK;
}

axs_status axs_update_vorticity_from_solid_stress_periodic_f64(
    double* vorticity, double* tau_z, double* tau_r, double* tau11, double* tau12, double* tau22,
    const double* R, int nr, int nz, double dt, double dx, int ghost, void* stream) {
  AXS_CHECK_ARG(vorticity && tau_z && tau_r && tau11 && tau12 && tau22 && R, "null pointer");
  AXS_CHECK_ARG(nr >= 3 && ghost > 0 && nz >= 4 * ghost, "need nr >= 3, nz >= 4 ghost");
  hipStream_t s = as_stream(stream);
  for (double* t : {tau11, tau12, tau22}) AXS_HIP(launch_ghost_comm(t, nr, nz, ghost, 0, 0.0, dx, s));
  div_tau_stage1_kernel<double, RFull<double>><<<dim3(ceil_div(nz - 2, 256), nr - 1), 256, 0, s>>>(
      tau_z, tau_r, tau11, tau12, tau22, RFull<double>{R, nz}, nr, nz, 2.0 * dx);
  AXS_LAUNCH_CHECK("div_tau_stage1");
  AXS_HIP(launch_ghost_comm(tau_r, nr, nz, ghost, 0, 0.0, dx, s));
  AXS_HIP(launch_ghost_comm(tau_z, nr, nz, ghost, 0, 0.0, dx, s));
  div_tau_stage2_kernel<double><<<dim3(ceil_div(nz - 2, 256), nr - 2), 256, 0, s>>>(
      vorticity, tau_z, tau_r, nr, nz, Scalar<double>{dt, nullptr}, 2.0 * dx);
  AXS_LAUNCH_CHECK("div_tau_stage2");
  return AXS_OK;
}

}  // extern "C"
