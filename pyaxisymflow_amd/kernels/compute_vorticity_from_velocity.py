"""Drop-in for pyaxisymflow/kernels/compute_vorticity_from_velocity.py:4-13."""
from pyaxisymflow_amd import _lib


def compute_vorticity_from_velocity_unb(vort, u_z, u_r, dx):
    """Interior curl of (u_z, u_r); border cells of vort are left untouched
    (compute_vorticity_from_velocity.py:4-13)."""
    st = _lib.Staged()
    nr, nz = _lib.same_shape(vort, u_z, u_r)
    _lib.call(
        "axs_compute_vorticity_from_velocity_unb_f64",
        st.ptr(vort, "vort", out=True),
        st.ptr(u_z, "u_z"),
        st.ptr(u_r, "u_r"),
        nr, nz, float(dx), _lib.stream_handle(),
    )
    st.finish()


def compute_vorticity_from_velocity_periodic(vort, u_z, u_r, dx, per_communicator):
    """compute_vorticity_from_velocity.py:16-28: the ghost exchange of u_r and
    u_z, then the interior stencil."""
    per_communicator(u_r)
    per_communicator(u_z)
    compute_vorticity_from_velocity_unb(vort, u_z, u_r, dx)
