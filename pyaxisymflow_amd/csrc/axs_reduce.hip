// Global reductions of the timestep: the Brinkman force / drag sum (row A9)
// and the CFL velocity maximum (row A10).  Two deterministic stages: a fixed
// grid of blocks writes one partial each, then one block combines them in a
// fixed order, so repeated runs give identical bits.
#include "axs_reduce.hpp"

#include <algorithm>

namespace axs {

int reduce_blocks(int64_t n) {
  int64_t b = (n + kReduceThreads * 4 - 1) / (kReduceThreads * 4);
  if (b < 1) b = 1;
  if (b > kReduceMaxBlocks) b = kReduceMaxBlocks;
  return static_cast<int>(b);
}

template <typename Op>
__device__ __forceinline__ double block_reduce(double v, Op op) {
  __shared__ double sm[kReduceThreads];
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int off = kReduceThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sm[threadIdx.x] = op(sm[threadIdx.x], sm[threadIdx.x + off]);
    __syncthreads();
  }
  return sm[0];
}

struct SumOp {
  __device__ double operator()(double a, double b) const { return a + b; }
  static constexpr double identity = 0.0;
};
struct MaxOp {
  __device__ double operator()(double a, double b) const { return a > b ? a : b; }
  static constexpr double identity = -1.0;  // every value reduced is >= 0
};

// sum over cells of (R * chi) * (u_z - U)   (compute_forces.py:14)
template <typename T, class RA>
__global__ void __launch_bounds__(kReduceThreads) force_partial_kernel(
    RA R, const T* __restrict__ chi, const T* __restrict__ uz, Scalar<double> U, int nz,
    int64_t n, double* __restrict__ partials) {
  const double Uv = U.get();
  double acc = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * kReduceThreads + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * kReduceThreads) {
    const int i = static_cast<int>(k / nz), j = static_cast<int>(k - (int64_t)i * nz);
    acc += (double(R.at(i, j)) * double(chi[k])) * (double(uz[k]) - Uv);
  }
  const double r = block_reduce(acc, SumOp{});
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// max over cells of |u_z| + |u_r|   (flow_past_sphere.py:152)
template <typename T>
__global__ void __launch_bounds__(kReduceThreads) maxvel_partial_kernel(
    const T* __restrict__ uz, const T* __restrict__ ur, int64_t n, double* __restrict__ partials) {
  double acc = -1.0;
  for (int64_t k = (int64_t)blockIdx.x * kReduceThreads + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * kReduceThreads) {
    const double v = double(fabs(uz[k]) + fabs(ur[k]));
    acc = v > acc ? v : acc;
  }
  const double r = block_reduce(acc, MaxOp{});
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

template <class Op>
__global__ void __launch_bounds__(kReduceThreads) finalize_kernel(const double* __restrict__ partials,
                                                                  int count, double* __restrict__ out,
                                                                  Op op) {
  double acc = Op::identity;
  for (int q = threadIdx.x; q < count; q += kReduceThreads) acc = op(acc, partials[q]);
  const double r = block_reduce(acc, op);
  if (threadIdx.x == 0) out[0] = r;
}

template <typename T, class RA>
hipError_t launch_force_sum(RA R, const T* chi, const T* uz, Scalar<double> U, int nz, int64_t n,
                            double* partials, double* out, hipStream_t s) {
  const int nb = reduce_blocks(n);
  force_partial_kernel<T, RA><<<nb, kReduceThreads, 0, s>>>(R, chi, uz, U, nz, n, partials);
  finalize_kernel<SumOp><<<1, kReduceThreads, 0, s>>>(partials, nb, out, SumOp{});
  return hipGetLastError();
}

template <typename T>
hipError_t launch_max_velocity(const T* uz, const T* ur, int64_t n, double* partials,
                               double* out, hipStream_t s) {
  const int nb = reduce_blocks(n);
  maxvel_partial_kernel<T><<<nb, kReduceThreads, 0, s>>>(uz, ur, n, partials);
  finalize_kernel<MaxOp><<<1, kReduceThreads, 0, s>>>(partials, nb, out, MaxOp{});
  return hipGetLastError();
}

template hipError_t launch_force_sum<double, RFull<double>>(RFull<double>, const double*,
                                                            const double*, Scalar<double>, int,
                                                            int64_t, double*, double*,
                                                            hipStream_t);
template hipError_t launch_force_sum<double, RRow<double>>(RRow<double>, const double*,
                                                           const double*, Scalar<double>, int,
                                                           int64_t, double*, double*,
                                                           hipStream_t);
template hipError_t launch_max_velocity<double>(const double*, const double*, int64_t, double*,
                                                double*, hipStream_t);

// Plain streaming copy (one read, one write per element): the reference
// point for achievable HBM bandwidth and for calibrating byte counters.
__global__ void __launch_bounds__(256) copy_f64_kernel(double* __restrict__ dst,
                                                       const double* __restrict__ src,
                                                       int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * 256 * 2;
  for (int64_t k = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 2; k < n; k += stride) {
    if (k + 1 < n) {
      const double2 v = *reinterpret_cast<const double2*>(src + k);
      *reinterpret_cast<double2*>(dst + k) = v;
    } else {
      dst[k] = src[k];
    }
  }
}

}  // namespace axs

using namespace axs;

extern "C" {

int64_t axs_reduce_partials_size(int64_t n) { return reduce_blocks(n); }

axs_status axs_device_copy_f64(double* dst, const double* src, int64_t n, void* stream) {
  AXS_CHECK_ARG(dst && src, "null pointer");
  AXS_CHECK_ARG(n >= 0, "negative size");
  AXS_CHECK_ARG((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) % 16 == 0,
                "pointers must be 16-byte aligned");
  if (n == 0) return AXS_OK;
  const int64_t blocks = std::min<int64_t>((n + 511) / 512, 256 * 32);
  copy_f64_kernel<<<(unsigned)blocks, 256, 0, as_stream(stream)>>>(dst, src, n);
  AXS_LAUNCH_CHECK("copy_f64");
  return AXS_OK;
}

axs_status axs_penalization_force_sum_f64(const double* R, const double* chi, const double* u_z,
                                          double U, int64_t n, double* partials,
                                          double* out_sum, void* stream) {
  AXS_CHECK_ARG(R && chi && u_z && partials && out_sum, "null pointer");
  AXS_CHECK_ARG(n >= 1 && n <= (int64_t)1 << 40, "bad size");
  // R is read as a flat array; nz only splits the flat index, any divisor works.
  AXS_HIP(launch_force_sum<double, RFull<double>>(RFull<double>{R, 1}, chi, u_z,
                                                  Scalar<double>{U, nullptr}, 1, n, partials,
                                                  out_sum, as_stream(stream)));
  return AXS_OK;
}

axs_status axs_max_abs_velocity_sum_f64(const double* u_z, const double* u_r, int64_t n,
                                        double* partials, double* out_max, void* stream) {
  AXS_CHECK_ARG(u_z && u_r && partials && out_max, "null pointer");
  AXS_CHECK_ARG(n >= 1, "bad size");
  AXS_HIP(launch_max_velocity<double>(u_z, u_r, n, partials, out_max, as_stream(stream)));
  return AXS_OK;
}

}  // extern "C"
