// Allocation-free variant of the penalisation force sum (row A9,
// kernels/compute_forces.py:14): the same two-stage deterministic
// reduction as axs_penalization_force_sum_f64, with its partial sums in a
// file-scope device array instead of a caller workspace, so that callers
// that cannot allocate device memory (code under stream capture, tight
// serving loops, harnesses that forbid hipMalloc) can take the sum.
//
// Each GPU holds its own copy of `scratch` for the life of the process.
// Host code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address): the device address is looked up once
// per device with hipGetSymbolAddress, with that device current, and kept
// in scratch_ptr[device].
#include <mutex>

#include "axs_reduce.hpp"

namespace axs {

// One partial per first-stage block; reduce_blocks() never asks for more
// than kReduceMaxBlocks (a grid-stride loop covers any n), so the array
// bounds every call: 1024 doubles = 8 KiB per GPU.
constexpr int kNoAllocPartials = kReduceMaxBlocks;
static __device__ __align__(256) double scratch[kNoAllocPartials];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex scratch_mu;
static double* scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of this device's scratch array (looked up on the first
// call on each device, with that device current).
static hipError_t scratch_address(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace axs

using namespace axs;

extern "C" {

axs_status axs_penalization_force_sum_noalloc_f64(const double* R, const double* chi,
                                                  const double* u_z, double U, int64_t n,
                                                  double* out_sum, void* stream) {
  AXS_CHECK_ARG(R && chi && u_z && out_sum, "null pointer");
  AXS_CHECK_ARG(n >= 1, "bad size");
  AXS_CHECK_ARG(n <= AXS_NOALLOC_MAX_CELLS,
                "n exceeds the allocation-free force sum's maximum of 2^40 cells");
  // the first stage writes every partial the second stage reads, so the
  // array needs no clearing between calls
  AXS_CHECK_ARG(reduce_blocks(n) <= kNoAllocPartials,
                "the reduction needs more partial sums than the scratch array holds");
  double* partials = nullptr;
  AXS_HIP(scratch_address(&partials));
  AXS_HIP(launch_force_sum<double, RFull<double>>(RFull<double>{R, 1}, chi, u_z,
                                                  Scalar<double>{U, nullptr}, 1, n, partials,
                                                  out_sum, as_stream(stream)));
  return AXS_OK;
}

}  // extern "C"
