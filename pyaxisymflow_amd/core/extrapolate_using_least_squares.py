"""Drop-in for the pybind11 module pyaxisymflow.core.extrapolate_using_least_squares
(core/extrapolate_using_least_squares_bind.cpp:11-100)."""
import ctypes

import numpy as np

from pyaxisymflow_amd import _lib

# Every sweep extrapolates one more front of cells; the band the drivers ask
# for is a few cells wide, so this bound is never reached in practice.
MAX_SWEEPS = 4096


def extrapolate_using_least_squares_till_first_order(
    current_flag, target_flag, eta_x, eta_y, grid_x, grid_y
):
    """Linear least-squares extrapolation of (eta_x, eta_y) from cells with
    current_flag set to every cell with target_flag set, front by front
    (core/src/extrapolate_using_least_squares.hpp:142-467).  current_flag,
    eta_x and eta_y are updated in place; flags are int16 and fields float64
    as the pybind11 binding requires (TypeError otherwise).  Returns the
    number of sweeps."""
    st = _lib.Staged()
    n0, n1 = _lib.same_shape(current_flag, target_flag, eta_x, eta_y)
    for g, n, name in ((grid_x, n1, "grid_x"), (grid_y, n0, "grid_y")):
        if len(g.shape) != 1 or g.shape[0] < n:
            raise ValueError(f"{name} needs at least {n} entries")
    done = ctypes.c_int(0)
    _lib.call(
        "axs_extrapolate_lsq_first_order_f64",
        st.ptr(current_flag, "current_flag", dtype=np.int16, out=True),
        st.ptr(target_flag, "target_flag", dtype=np.int16),
        st.ptr(eta_x, "eta_x", out=True), st.ptr(eta_y, "eta_y", out=True),
        st.ptr(grid_x, "grid_x"), st.ptr(grid_y, "grid_y"),
        n0, n1, MAX_SWEEPS, ctypes.byref(done), _lib.stream_handle(),
    )
    st.finish()
    return done.value
