// Deterministic two-stage reductions shared by the drop-in entry points and
// the stepper.
#pragma once

#include "axs_common.hpp"

namespace axs {

constexpr int kReduceThreads = 256;
constexpr int kReduceMaxBlocks = 1024;

// Number of first-stage blocks (= partials needed) for n cells.
int reduce_blocks(int64_t n);

template <typename T, class RA>
hipError_t launch_force_sum(RA R, const T* chi, const T* uz, Scalar<double> U, int nz, int64_t n,
                            double* partials, double* out, hipStream_t s);

template <typename T>
hipError_t launch_max_velocity(const T* uz, const T* ur, int64_t n, double* partials,
                               double* out, hipStream_t s);

}  // namespace axs
