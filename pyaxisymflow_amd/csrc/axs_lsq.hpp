// Least-squares extrapolation sweep shared by the drop-in entry point and
// the soft-body stepper.
#pragma once

#include "axs_common.hpp"

namespace axs {

struct LsqWork {
  double* nx;          // fitted eta_x of this sweep, (n0, n1)
  double* ny;          // fitted eta_y of this sweep, (n0, n1)
  int8_t* pend;        // 1 where a cell was fitted this sweep
  unsigned* counters;  // cells fitted per sweep (one slot per sweep)
};

// One sweep (fit + commit).  Sweep `sweep` exits at once when
// counters[sweep - 1] is zero, so many sweeps can be queued without a sync.
hipError_t launch_lsq_sweep(int16_t* cur, const int16_t* tgt, double* ex, double* ey,
                            const double* gx, const double* gy, int n0, int n1, LsqWork w,
                            int sweep, hipStream_t s);
hipError_t launch_fill_u32(unsigned* p, int n, unsigned v, hipStream_t s);

}  // namespace axs
