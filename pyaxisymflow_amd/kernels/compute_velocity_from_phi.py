"""Drop-in for pyaxisymflow/kernels/compute_velocity_from_phi.py:4-17."""
from pyaxisymflow_amd import _lib


def compute_velocity_from_phi_unb(u_z, u_r, phi, dx):
    """u_z = dphi/dz, u_r = dphi/dr (2nd-order central, one-sided at the
    edges); compute_velocity_from_phi.py:4-17."""
    st = _lib.Staged()
    nr, nz = _lib.same_shape(u_z, u_r, phi)
    _lib.call(
        "axs_velocity_from_phi_f64",
        st.ptr(u_z, "u_z", out=True, copy_in=False),
        st.ptr(u_r, "u_r", out=True, copy_in=False),
        st.ptr(phi, "phi"),
        nr, nz, float(dx), _lib.stream_handle(),
    )
    st.finish()
