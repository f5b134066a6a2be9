"""Drop-in for pyaxisymflow/kernels/compute_velocity_from_psi.py:4-17."""
from pyaxisymflow_amd import _lib


def compute_velocity_from_psi_unb(u_z, u_r, psi, R, dx):
    """u_z = dpsi/dr / r, u_r = -dpsi/dz / r (2nd-order central, one-sided
    at the edges); compute_velocity_from_psi.py:4-17."""
    st = _lib.Staged()
    nr, nz = _lib.same_shape(u_z, u_r, psi, R)
    _lib.call(
        "axs_compute_velocity_from_psi_unb_f64",
        st.ptr(u_z, "u_z", out=True, copy_in=False),
        st.ptr(u_r, "u_r", out=True, copy_in=False),
        st.ptr(psi, "psi"),
        st.ptr(R, "R"),
        nr, nz, float(dx), _lib.stream_handle(),
    )
    st.finish()


def compute_velocity_from_psi_periodic(u_z, u_r, psi, R, dx, per_communicator):
    """compute_velocity_from_psi.py:20-35: the ghost exchange of psi, then the
    unbounded stencil."""
    per_communicator(psi)
    compute_velocity_from_psi_unb(u_z, u_r, psi, R, dx)
