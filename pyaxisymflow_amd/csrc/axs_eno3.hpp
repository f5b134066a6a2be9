// ENO3 advection launchers shared by the drop-in entry points and the stepper.
#pragma once

#include "axs_common.hpp"

namespace axs {

// out = f + conservative ENO3 flux; f has axis parity `parity` (+1 even, -1 odd).
template <typename T>
hipError_t launch_eno3_conservative(const T* f, const T* uz, const T* ur, T* out, int nr, int nz,
                                    Scalar<T> dt_by_dx, T parity, hipStream_t s);

// (o1, o2) = (e1, e2) + non-conservative ENO3 flux; e1 even, e2 odd.
template <typename T>
hipError_t launch_eno3_refmap(const T* e1, const T* e2, const T* uz, const T* ur, T* o1, T* o2,
                              int nr, int nz, Scalar<T> dt_by_dx, hipStream_t s);

}  // namespace axs
