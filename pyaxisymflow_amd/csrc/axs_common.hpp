// Shared definitions for the axisymmetric vorticity-streamfunction kernels.
// Target: gfx950 (CDNA4, 64-wide wavefronts).  Fields are row-major (nr, nz):
// row i is radial index r_i, column j is axial index z_j (contiguous).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/axisym.h"

namespace axs {

// Thread-local last-error string reported by axs_last_error().
void set_error(const std::string& msg);
void clear_error();

#define AXS_CHECK_ARG(cond, msg)            \
  do {                                      \
    if (!(cond)) {                          \
      ::axs::set_error(std::string(msg));   \
      return AXS_ERR_INVALID;               \
    }                                       \
  } while (0)

#define AXS_HIP(...)                                                         \
  do {                                                                       \
    hipError_t _e = (__VA_ARGS__);                                           \
    if (_e != hipSuccess) {                                                  \
      ::axs::set_error(std::string(#__VA_ARGS__) + ": " + hipGetErrorString(_e)); \
      return AXS_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)

// Checks the launch just issued; returns from the calling C-ABI function.
#define AXS_LAUNCH_CHECK(name) AXS_HIP(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned ceil_div(int64_t a, int64_t b) { return static_cast<unsigned>((a + b - 1) / b); }

// A scalar that is either a host value or read from device memory (so a
// whole timestep can run with dt resident on the device).
template <typename T>
struct Scalar {
  T host;
  const double* dev;
  __device__ __forceinline__ T get() const { return dev ? static_cast<T>(*dev) : host; }
};

// Radial coordinate: either a full (nr, nz) array as the reference passes R,
// or one value per row (the grid is row-constant in r).
template <typename T>
struct RFull {
  const T* R;
  int nz;
  __device__ __forceinline__ T at(int i, int j) const { return R[(int64_t)i * nz + j]; }
};
template <typename T>
struct RRow {
  const double* r;
  __device__ __forceinline__ T at(int i, int) const { return static_cast<T>(r[i]); }
};

// x / d for a divisor constant per launch or per row, given y = RN(1 / d).
// FP64 (the reference's arithmetic, matched bit for bit): rounded to
// nearest.  q = RN(x y) is within one ulp of x / d, the residual x - q d is
// exact under fma, and q + RN(r y) is then the correctly rounded quotient
// (Markstein's correction step, P. Markstein, "Computation of elementary
// functions on the IBM RISC System/6000 processor", IBM J. Res. Dev.
// 34(1), 1990).  Holds while neither the quotient nor the residual
// overflows or underflows; a zero quotient comes back as +0.  Three FP ops
// instead of the IEEE division sequence.  FP32 (the precision and batched
// forms, held to 1e-4 of FP64, not to bits): x y, within an ulp.
template <typename T>
__device__ __forceinline__ T div_by(T x, T d, T y) {
  if constexpr (sizeof(T) == 4) {
    return x * y;
  } else {
    const T q = x * y;
    const T r = fma(-q, d, x);
    return fma(r, y, q);
  }
}

// Maximum of two values that are non-negative or NaN, NaN winning, as
// numpy's amax propagates NaN (fmax drops it): with the sign bit cleared,
// non-negative doubles order as their bit patterns do and every NaN lies
// above +inf.  Equal to fmax on non-negative numbers, bit for bit.
__device__ __forceinline__ double max_nan(double a, double b) {
  const long long x = __double_as_longlong(a) & 0x7fffffffffffffffLL;
  const long long y = __double_as_longlong(b) & 0x7fffffffffffffffLL;
  return __longlong_as_double(x > y ? x : y);
}

// Appends value to list for the lanes with take set: one atomic per wave
// for all of them (a counter that every lane of a large grid bumps alone
// serialises the whole pass), lanes in lane order after the wave's base.
__device__ __forceinline__ void list_append(bool take, int value, int* __restrict__ list,
                                            unsigned* __restrict__ count) {
  const unsigned long long mask = __ballot(take);
  if (mask == 0ull) return;
  const int lane = __lane_id();
  const int leader = __ffsll(static_cast<long long>(mask)) - 1;
  unsigned base = 0u;
  if (lane == leader) base = atomicAdd(count, static_cast<unsigned>(__popcll(mask)));
  base = __shfl(base, leader);
  if (take) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = value;
}

}  // namespace axs
