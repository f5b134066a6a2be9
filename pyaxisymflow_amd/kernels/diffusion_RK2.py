"""Drop-in for pyaxisymflow/kernels/diffusion_RK2.py:4-44."""
from pyaxisymflow_amd import _lib


def diffusion_RK2_unb(vorticity, temp_vorticity, R, nu, dt, dx):
    """Midpoint RK2 of nu (lap w + w_r/r - w/r^2) on the interior
    (diffusion_RK2.py:4-44); temp_vorticity holds the half step."""
    st = _lib.Staged()
    nr, nz = _lib.same_shape(vorticity, temp_vorticity, R)
    _lib.call(
        "axs_diffusion_rk2_unb_f64",
        st.ptr(vorticity, "vorticity", out=True),
        st.ptr(temp_vorticity, "temp_vorticity", out=True, copy_in=False),
        st.ptr(R, "R"),
        nr, nz, float(nu), float(dt), float(dx), _lib.stream_handle(),
    )
    st.finish()


def diffusion_RK2_periodic(vorticity, temp_vorticity, R, nu, dt, dx, per_communicator):
    """diffusion_RK2.py:47-90: the ghost exchange before each RK2 stage
    (per_communicator from gen_periodic_boundary_ghost_comm)."""
    from pyaxisymflow_amd.kernels.periodic_boundary_ghost_comm import plain_ghost_size

    g = plain_ghost_size(per_communicator, "diffusion_RK2_periodic")
    st = _lib.Staged()
    nr, nz = _lib.same_shape(vorticity, temp_vorticity, R)
    _lib.call(
        "axs_diffusion_rk2_periodic_f64",
        st.ptr(vorticity, "vorticity", out=True),
        st.ptr(temp_vorticity, "temp_vorticity", out=True, copy_in=False),
        st.ptr(R, "R"),
        nr, nz, float(nu), float(dt), float(dx), g, _lib.stream_handle(),
    )
    st.finish()
