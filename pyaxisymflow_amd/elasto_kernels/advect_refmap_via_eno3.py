"""Drop-in for pyaxisymflow/elasto_kernels/advect_refmap_via_eno3.py:8-54."""
import ctypes

import numpy as np

from pyaxisymflow_amd import _lib


def gen_advect_refmap_via_eno3(
    dx, grid_size_r, grid_size_z, real_t=np.float64, num_threads=False
):
    """Returns advect_refmap_via_eno3(eta1, eta2, u_z, u_r, dt): one
    non-conservative ENO3 Euler step of the reference map, eta1 even and
    eta2 odd about the axis (advect_refmap_via_eno3.py:8-54)."""
    if real_t not in (np.float64, float):
        raise NotImplementedError("the drop-in ENO3 step is FP64 (the reference's is too)")
    holder = {}

    def advect_refmap_via_eno3(eta1, eta2, u_z, u_r, dt):
        st = _lib.Staged()
        nr, nz = _lib.same_shape(eta1, eta2, u_z, u_r)
        if (nr, nz) != (grid_size_r, grid_size_z):
            raise ValueError(f"field shape {(nr, nz)} != {(grid_size_r, grid_size_z)}")
        p1 = st.ptr(eta1, "eta1", out=True)
        p2 = st.ptr(eta2, "eta2", out=True)
        buf = holder.get("scratch")
        if buf is None or buf.device != st.dev:
            buf = holder["scratch"] = st.scratch((2, nr, nz))
        _lib.call(
            "axs_advect_refmap_eno3_f64",
            p1, p2, st.ptr(u_z, "u_z"), st.ptr(u_r, "u_r"), ctypes.c_void_p(buf.data_ptr()),
            nr, nz, float(dt), float(dx), _lib.stream_handle(),
        )
        st.finish()

    return advect_refmap_via_eno3


def gen_advect_refmap_via_eno3_periodic(
    dx, grid_size_r, grid_size_z, per_communicator1, per_communicator2, real_t=np.float64,
    num_threads=False,
):
    """advect_refmap_via_eno3.py:57-116: the ghost exchange of u_z, u_r
    (per_communicator1), eta1 (per_communicator2, the shifted exchange) and
    eta2 (per_communicator1), then the same ENO3 steps."""
    step = gen_advect_refmap_via_eno3(dx, grid_size_r, grid_size_z, real_t, num_threads)

    def advect_refmap_via_eno3_periodic(eta1, eta2, u_z, u_r, dt):
        per_communicator1(u_z)
        per_communicator1(u_r)
        per_communicator2(eta1)
        per_communicator1(eta2)
        step(eta1, eta2, u_z, u_r, dt)

    return advect_refmap_via_eno3_periodic
