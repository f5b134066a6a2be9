"""Drop-in for pyaxisymflow/kernels/advect_vorticity_via_eno3.py:8-44."""
import ctypes

import numpy as np

from pyaxisymflow_amd import _lib


def gen_advect_vorticity_via_eno3(
    dx, grid_size_r, grid_size_z, real_t=np.float64, num_threads=False
):
    """Returns advect_vorticity_via_eno3(vorticity, u_z, u_r, dt): one
    conservative ENO3 Euler step with omega odd and u_r odd about the axis
    (advect_vorticity_via_eno3.py:8-44).  `num_threads` is accepted for
    signature compatibility; the closure owns one (nr, nz) device buffer
    (the reference owns three doubled ones)."""
    if real_t not in (np.float64, float):
        raise NotImplementedError("the drop-in ENO3 step is FP64 (the reference's is too)")
    holder = {}

    def advect_vorticity_via_eno3(vorticity, u_z, u_r, dt):
        st = _lib.Staged()
        nr, nz = _lib.same_shape(vorticity, u_z, u_r)
        if (nr, nz) != (grid_size_r, grid_size_z):
            raise ValueError(f"field shape {(nr, nz)} != {(grid_size_r, grid_size_z)}")
        pw = st.ptr(vorticity, "vorticity", out=True)
        buf = holder.get("scratch")
        if buf is None or buf.device != st.dev:
            buf = holder["scratch"] = st.scratch((nr, nz))
        _lib.call(
            "axs_advect_vorticity_eno3_f64",
            pw, st.ptr(u_z, "u_z"), st.ptr(u_r, "u_r"), ctypes.c_void_p(buf.data_ptr()),
            nr, nz, float(dt), float(dx), _lib.stream_handle(),
        )
        st.finish()

    return advect_vorticity_via_eno3


def gen_advect_vorticity_via_eno3_periodic(
    dx, grid_size_r, grid_size_z, per_communicator, real_t=np.float64, num_threads=False
):
    """advect_vorticity_via_eno3.py:47-91: the ghost exchange of u_z, u_r and
    the vorticity, then the same ENO3 step."""
    step = gen_advect_vorticity_via_eno3(dx, grid_size_r, grid_size_z, real_t, num_threads)

    def advect_vorticity_via_eno3_periodic(vorticity, u_z, u_r, dt):
        per_communicator(u_z)
        per_communicator(u_r)
        per_communicator(vorticity)
        step(vorticity, u_z, u_r, dt)

    return advect_vorticity_via_eno3_periodic
