// ENO3 Euler advection step (rows A7 and A13).
//
// The reference mirrors every field across the axis into a (2nr, nz) buffer
// (kernels/advect_vorticity_via_eno3.py:28-33: u_z and eta1 even, u_r,
// omega and eta2 odd), runs pystencils kernels over rows/cols [2, n-2) of
// that buffer (ghost layers = largest stencil offset = 2), and copies the
// physical half back.  Here the mirror is applied in the index: physical row
// m < 0 reads row -m-1 with the field's parity, so the doubled buffers never
// exist.  Updated cells: rows [0, nr-2), cols [2, nz-2); the rest keep their
// value (field + 0).
//
// Flux arithmetic follows pyst_kernels/advection_flux.py:32-44, 61-73,
// 88-100, 115-127 (conservative) and :189-202, 219-236, 251-264, 279-296
// (non-conservative), accumulated in the order x-front, x-back, y-front,
// y-back with inv_dx = -dt/dx (advection_timestep.py:45-54).
#include "axs_eno3.hpp"
#include "axs_softbatch.hpp"

namespace axs {

// Conservative face value between cells q (centre side) and its neighbour in
// the direction of `front`: pass the five (f*u) products and u at offsets
// -2..+2 around the centre.
template <typename T>
__device__ __forceinline__ T eno3_front(const T* f, const T* u) {
  // index 2 is the centre
  return (u[2] > -u[3])
             ? (T(1) / T(3)) * f[3] * u[3] + (T(5) / T(6)) * f[2] * u[2] - (T(1) / T(6)) * f[1] * u[1]
             : (T(1) / T(3)) * f[2] * u[2] + (T(5) / T(6)) * f[3] * u[3] - (T(1) / T(6)) * f[4] * u[4];
}
template <typename T>
__device__ __forceinline__ T eno3_back(const T* f, const T* u) {
  return (u[2] > -u[1])
             ? (T(1) / T(3)) * f[2] * u[2] + (T(5) / T(6)) * f[1] * u[1] - (T(1) / T(6)) * f[0] * u[0]
             : (T(1) / T(3)) * f[1] * u[1] + (T(5) / T(6)) * f[2] * u[2] - (T(1) / T(6)) * f[3] * u[3];
}
// Non-conservative face interpolants (no velocity inside).
template <typename T>
__device__ __forceinline__ T eno3_front_nc(const T* f, const T* u) {
  return (u[2] > -u[3]) ? (T(1) / T(3)) * f[3] + (T(5) / T(6)) * f[2] - (T(1) / T(6)) * f[1]
                        : (T(1) / T(3)) * f[2] + (T(5) / T(6)) * f[3] - (T(1) / T(6)) * f[4];
}
template <typename T>
__device__ __forceinline__ T eno3_back_nc(const T* f, const T* u) {
  return (u[2] > -u[1]) ? (T(1) / T(3)) * f[2] + (T(5) / T(6)) * f[1] - (T(1) / T(6)) * f[0]
                        : (T(1) / T(3)) * f[1] + (T(5) / T(6)) * f[2] - (T(1) / T(6)) * f[3];
}

// Row index and sign of physical row m (m may be -1 or -2: mirrored).
__device__ __forceinline__ int mirror_row(int m, int* flip) {
  if (m < 0) {
    *flip = 1;
    return -m - 1;
  }
  *flip = 0;
  return m;
}

// Stencil loads, converted to the arithmetic type T (S = T but for the
// FP32 velocity of the FP32-bulk soft batch's FP64 reference-map step).
template <typename S, typename T>
__device__ __forceinline__ void load_column5(const S* __restrict__ a, int i, int j, int nz,
                                             T odd_sign, T* out) {
#pragma unroll
  for (int o = -2; o <= 2; ++o) {
    int flip;
    const int m = mirror_row(i + o, &flip);
    const T v = T(a[(int64_t)m * nz + j]);
    out[o + 2] = flip ? odd_sign * v : v;
  }
}

template <typename S, typename T>
__device__ __forceinline__ void load_row5(const S* __restrict__ a, int i, int j, int nz,
                                          T* out) {
  const S* row = a + (int64_t)i * nz + j;
#pragma unroll
  for (int o = -2; o <= 2; ++o) out[o + 2] = T(row[o]);
}

// One conservative ENO3 step at cell (i, j): out = f + flux(f, u), f with
// axis parity f_parity.
template <typename T>
__device__ __forceinline__ void eno3_conservative_at(const T* __restrict__ f,
                                                     const T* __restrict__ uz,
                                                     const T* __restrict__ ur,
                                                     T* __restrict__ out, int nr, int nz,
                                                     T dt_by_dx, T f_parity, int i, int j) {
  const int64_t c = (int64_t)i * nz + j;
  if (i >= nr - 2 || j < 2 || j >= nz - 2) {
    out[c] = f[c];
    return;
  }
  const T s = -dt_by_dx;
  T fx[5], ux[5], fy[5], uy[5];
  load_row5(f, i, j, nz, fx);
  load_row5(uz, i, j, nz, ux);
  load_column5(f, i, j, nz, f_parity, fy);
  load_column5(ur, i, j, nz, T(-1), uy);
  T flux = T(0) + s * eno3_front(fx, ux);
  flux = flux - s * eno3_back(fx, ux);
  flux = flux + s * eno3_front(fy, uy);
  flux = flux - s * eno3_back(fy, uy);
  out[c] = f[c] + flux;
}

template <typename T>
__global__ void eno3_conservative_kernel(const T* __restrict__ f, const T* __restrict__ uz,
                                         const T* __restrict__ ur, T* __restrict__ out, int nr,
                                         int nz, Scalar<T> dt_by_dx, T f_parity) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  eno3_conservative_at<T>(f, uz, ur, out, nr, nz, dt_by_dx.get(), f_parity, i, j);
}

// One non-conservative ENO3 step for two fields sharing the velocity at cell
// (i, j): eta1 (even) and eta2 (odd); the velocity stored in U.
template <typename T, typename U = T>
__device__ __forceinline__ void eno3_nonconservative2_at(
    const T* __restrict__ e1, const T* __restrict__ e2, const U* __restrict__ uz,
    const U* __restrict__ ur, T* __restrict__ o1, T* __restrict__ o2, int nr, int nz,
    T dt_by_dx, int i, int j) {
  const int64_t c = (int64_t)i * nz + j;
  if (i >= nr - 2 || j < 2 || j >= nz - 2) {
    o1[c] = e1[c];
    o2[c] = e2[c];
    return;
  }
  const T s = -dt_by_dx;
  T ux[5], uy[5];
  load_row5(uz, i, j, nz, ux);
  load_column5(ur, i, j, nz, T(-1), uy);
  const T uxc = ux[2], uyc = uy[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const T* f = q == 0 ? e1 : e2;
    T fx[5], fy[5];
    load_row5(f, i, j, nz, fx);
    load_column5(f, i, j, nz, q == 0 ? T(1) : T(-1), fy);
    T flux = T(0) + s * eno3_front_nc(fx, ux) * uxc;
    flux = flux - s * eno3_back_nc(fx, ux) * uxc;
    flux = flux + s * eno3_front_nc(fy, uy) * uyc;
    flux = flux - s * eno3_back_nc(fy, uy) * uyc;
    (q == 0 ? o1 : o2)[c] = f[c] + flux;
  }
}

template <typename T>
__global__ void eno3_nonconservative2_kernel(const T* __restrict__ e1, const T* __restrict__ e2,
                                             const T* __restrict__ uz, const T* __restrict__ ur,
                                             T* __restrict__ o1, T* __restrict__ o2, int nr,
                                             int nz, Scalar<T> dt_by_dx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  eno3_nonconservative2_at<T>(e1, e2, uz, ur, o1, o2, nr, nz, dt_by_dx.get(), i, j);
}

// Soft batches: one row segment of 256 cells of one case per workgroup.
// The reference map (FP64, with the batch's velocity): (eta1, eta2) ->
// (e1n, e2n).
template <typename B>
__global__ void soft_batch_refmap_kernel(const SoftCaseT<B>* __restrict__ cs,
                                         const uint32_t* __restrict__ blocks) {
  const SoftBlock b = soft_block(blocks);
  const SoftCaseT<B>& C = cs[b.c];
  // beyond the last zone the reference map is zero and stays so this step
  // (the pin skips the same rows, so e1n / e2n there are never read)
  if (b.j >= C.nz || !prev_zone_row(C, b.i)) return;
  eno3_nonconservative2_at<double, B>(C.eta1, C.eta2, C.u_z, C.u_r, C.e1n, C.e2n, C.nr, C.nz,
                                      C.scalars[kSsDtByDx], b.i, b.j);
}

// The vorticity: w0 (after the stress update) -> w1.
template <typename B>
__global__ void soft_batch_advect_kernel(const SoftCaseT<B>* __restrict__ cs,
                                         const uint32_t* __restrict__ blocks) {
  const SoftBlock b = soft_block(blocks);
  const SoftCaseT<B>& C = cs[b.c];
  if (b.j >= C.nz) return;
  eno3_conservative_at<B>(C.w0, C.u_z, C.u_r, C.w1, C.nr, C.nz, B(C.scalars[kSsDtByDx]), B(-1),
                          b.i, b.j);
}

template <typename B>
hipError_t launch_soft_batch_refmap(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                    hipStream_t s) {
  soft_batch_refmap_kernel<B><<<d.ncell_blocks, 256, 0, s>>>(cases, d.cell_blocks);
  return hipGetLastError();
}

template <typename B>
hipError_t launch_soft_batch_advect(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                    hipStream_t s) {
  soft_batch_advect_kernel<B><<<d.ncell_blocks, 256, 0, s>>>(cases, d.cell_blocks);
  return hipGetLastError();
}

template hipError_t launch_soft_batch_refmap<double>(const SoftCaseT<double>*,
                                                     const SoftBatchDims&, hipStream_t);
template hipError_t launch_soft_batch_refmap<float>(const SoftCaseT<float>*, const SoftBatchDims&,
                                                    hipStream_t);
template hipError_t launch_soft_batch_advect<double>(const SoftCaseT<double>*,
                                                     const SoftBatchDims&, hipStream_t);
template hipError_t launch_soft_batch_advect<float>(const SoftCaseT<float>*, const SoftBatchDims&,
                                                    hipStream_t);

template <typename T>
hipError_t launch_eno3_conservative(const T* f, const T* uz, const T* ur, T* out, int nr, int nz,
                                    Scalar<T> dt_by_dx, T parity, hipStream_t s) {
  dim3 grid(ceil_div(nz, 256), nr);
  eno3_conservative_kernel<T><<<grid, 256, 0, s>>>(f, uz, ur, out, nr, nz, dt_by_dx, parity);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_eno3_refmap(const T* e1, const T* e2, const T* uz, const T* ur, T* o1, T* o2,
                              int nr, int nz, Scalar<T> dt_by_dx, hipStream_t s) {
  dim3 grid(ceil_div(nz, 256), nr);
  eno3_nonconservative2_kernel<T><<<grid, 256, 0, s>>>(e1, e2, uz, ur, o1, o2, nr, nz, dt_by_dx);
  return hipGetLastError();
}

template hipError_t launch_eno3_conservative<double>(const double*, const double*, const double*,
                                                     double*, int, int, Scalar<double>, double,
                                                     hipStream_t);
template hipError_t launch_eno3_conservative<float>(const float*, const float*, const float*,
                                                    float*, int, int, Scalar<float>, float,
                                                    hipStream_t);
template hipError_t launch_eno3_refmap<float>(const float*, const float*, const float*,
                                              const float*, float*, float*, int, int,
                                              Scalar<float>, hipStream_t);
template hipError_t launch_eno3_refmap<double>(const double*, const double*, const double*,
                                               const double*, double*, double*, int, int,
                                               Scalar<double>, hipStream_t);

}  // namespace axs

using namespace axs;

extern "C" {

axs_status axs_advect_vorticity_eno3_f64(double* vorticity, const double* u_z,
                                         const double* u_r, double* scratch, int nr, int nz,
                                         double dt, double dx, void* stream) {
  AXS_CHECK_ARG(vorticity && u_z && u_r && scratch, "null pointer");
  AXS_CHECK_ARG(nr >= 3 && nz >= 5, "need nr >= 3 and nz >= 5");
  AXS_CHECK_ARG(scratch != vorticity, "scratch must not alias vorticity");
  hipStream_t s = as_stream(stream);
  AXS_HIP(launch_eno3_conservative<double>(vorticity, u_z, u_r, scratch, nr, nz,
                                           Scalar<double>{dt / dx, nullptr}, -1.0, s));
  AXS_HIP(hipMemcpyAsync(vorticity, scratch, sizeof(double) * (size_t)nr * nz,
                         hipMemcpyDeviceToDevice, s));
  return AXS_OK;
}

axs_status axs_advect_refmap_eno3_f64(double* eta1, double* eta2, const double* u_z,
                                      const double* u_r, double* scratch, int nr, int nz,
                                      double dt, double dx, void* stream) {
  AXS_CHECK_ARG(eta1 && eta2 && u_z && u_r && scratch, "null pointer");
  AXS_CHECK_ARG(nr >= 3 && nz >= 5, "need nr >= 3 and nz >= 5");
  hipStream_t s = as_stream(stream);
  const size_t n = (size_t)nr * nz;
  AXS_HIP(launch_eno3_refmap<double>(eta1, eta2, u_z, u_r, scratch, scratch + n, nr, nz,
                                     Scalar<double>{dt / dx, nullptr}, s));
  AXS_HIP(hipMemcpyAsync(eta1, scratch, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  AXS_HIP(hipMemcpyAsync(eta2, scratch + n, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  return AXS_OK;
}

}  // extern "C"
