"""Drop-in for pyaxisymflow/elasto_kernels/solid_sigma.py:4-29."""
from pyaxisymflow_amd import _lib


def solid_sigma(
    sigma_s_11, sigma_s_12, sigma_s_22, G, dx, eta_1, eta_2, eta_1z, eta_1r, eta_2z, eta_2r
):
    """Elastic stress from reference-map gradients (solid_sigma.py:4-29).
    The gradient arrays are in/out scratch: entries the reference never
    writes keep their previous values and still enter sigma."""
    st = _lib.Staged()
    nr, nz = _lib.same_shape(sigma_s_11, sigma_s_12, sigma_s_22, eta_1, eta_2,
                             eta_1z, eta_1r, eta_2z, eta_2r)
    _lib.call(
        "axs_solid_sigma_f64",
        st.ptr(sigma_s_11, "sigma_s_11", out=True, copy_in=False),
        st.ptr(sigma_s_12, "sigma_s_12", out=True, copy_in=False),
        st.ptr(sigma_s_22, "sigma_s_22", out=True, copy_in=False),
        float(G), float(dx),
        st.ptr(eta_1, "eta_1"), st.ptr(eta_2, "eta_2"),
        st.ptr(eta_1z, "eta_1z", out=True), st.ptr(eta_1r, "eta_1r", out=True),
        st.ptr(eta_2z, "eta_2z", out=True), st.ptr(eta_2r, "eta_2r", out=True),
        nr, nz, _lib.stream_handle(),
    )
    st.finish()


def solid_sigma_periodic(
    sigma_s_11, sigma_s_12, sigma_s_22, G, dx, eta_1, eta_2, eta_1z, eta_1r, eta_2z, eta_2r,
    per_communicator1, per_communicator2,
):
    """solid_sigma.py:32-60: the ghost exchange of eta_1 (per_communicator2)
    and eta_2 (per_communicator1), then the same stress."""
    per_communicator2(eta_1)
    per_communicator1(eta_2)
    solid_sigma(sigma_s_11, sigma_s_12, sigma_s_22, G, dx, eta_1, eta_2, eta_1z, eta_1r, eta_2z,
                eta_2r)
