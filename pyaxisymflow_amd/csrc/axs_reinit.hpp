// Narrow-band signed-distance reinitialisation of a level set (the call
// skfmm.distance(phi, dx=dx, narrow=band) at
// examples/SoftSphereInTaylorGreenVortex/soft_sphere_in_taylor_green_vortex.py:159-162).
#pragma once

#include "axs_common.hpp"

namespace axs {

struct ReinitWork {
  double* d0;       // distance iterate (ping)
  double* d1;       // distance iterate (pong)
  double* dist;     // interface distances of the frozen cells
  uint8_t* frozen;  // 1 on interface cells
  uint8_t* ch0;     // 1 where the last sweep changed the cell (ping)
  uint8_t* ch1;     // (pong)
  // sweep work lists over the 8 x 32 tiles (T = reinit_tile_total):
  unsigned* tiles;  // 2 T flags, zero before the first use (each reinit
                    // leaves them zero)
  unsigned* list;   // 2 T entries
  unsigned* cnt;    // 3 counters
};

// Tiles of an nr x nz grid.
int reinit_tile_total(int nr, int nz);

// Sweeps the reference count of Jacobi iterations for a band.
int reinit_iterations(double narrow, double dx);

// out = signed distance where |d| <= narrow, phi elsewhere (masked = 1
// there, when masked is not null).  `iterations` Jacobi sweeps of the
// upwind eikonal update.  phi and out may alias.
hipError_t launch_reinit(const double* phi, double* out, uint8_t* masked, int nr, int nz,
                         double dx, double narrow, int iterations, const ReinitWork& w,
                         hipStream_t s);

}  // namespace axs
