"""Drop-in for pyaxisymflow/elasto_kernels/div_tau.py:4-34."""
from pyaxisymflow_amd import _lib


def update_vorticity_from_solid_stress(vorticity, tau_z, tau_r, tau11, tau12, tau22, R, dt, dx):
    """tau = div(sigma) with the axisymmetric hoop terms, then
    omega += dt curl(tau) on the interior (div_tau.py:4-34)."""
    st = _lib.Staged()
    nr, nz = _lib.same_shape(vorticity, tau_z, tau_r, tau11, tau12, tau22, R)
    _lib.call(
        "axs_update_vorticity_from_solid_stress_f64",
        st.ptr(vorticity, "vorticity", out=True),
        st.ptr(tau_z, "tau_z", out=True), st.ptr(tau_r, "tau_r", out=True),
        st.ptr(tau11, "tau11"), st.ptr(tau12, "tau12"), st.ptr(tau22, "tau22"),
        st.ptr(R, "R"), nr, nz, float(dt), float(dx), _lib.stream_handle(),
    )
    st.finish()


def update_vorticity_from_solid_stress_periodic(
    vorticity, tau_z, tau_r, tau11, tau12, tau22, R, dt, dx, per_communicator
):
    """div_tau.py:37-75: the ghost exchange of tau11, tau12, tau22 (in place),
    tau = div(sigma), the exchange of tau_r and tau_z, then the vorticity
    update (per_communicator from gen_periodic_boundary_ghost_comm)."""
    from pyaxisymflow_amd.kernels.periodic_boundary_ghost_comm import plain_ghost_size

    g = plain_ghost_size(per_communicator, "update_vorticity_from_solid_stress_periodic")
    st = _lib.Staged()
    nr, nz = _lib.same_shape(vorticity, tau_z, tau_r, tau11, tau12, tau22, R)
    _lib.call(
        "axs_update_vorticity_from_solid_stress_periodic_f64",
        st.ptr(vorticity, "vorticity", out=True),
        st.ptr(tau_z, "tau_z", out=True), st.ptr(tau_r, "tau_r", out=True),
        st.ptr(tau11, "tau11", out=True), st.ptr(tau12, "tau12", out=True),
        st.ptr(tau22, "tau22", out=True),
        st.ptr(R, "R"), nr, nz, float(dt), float(dx), g, _lib.stream_handle(),
    )
    st.finish()
