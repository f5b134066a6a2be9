"""Drop-in for pyaxisymflow/kernels/compute_forces.py:4-17."""
import ctypes

from pyaxisymflow_amd import _lib


def penalization_force_sum(R, part_char_func, u_z, U_z_cm_part):
    """sum(R * chi * (u_z - U)) in FP64 on the device (returns a float)."""
    st = _lib.Staged()
    shape = _lib.same_shape(R, part_char_func, u_z)
    n = shape[0] * shape[1]
    nparts = _lib.load().axs_reduce_partials_size(n)
    partials = st.scratch((nparts,))
    out = st.scratch((1,))
    _lib.call(
        "axs_penalization_force_sum_f64",
        st.ptr(R, "R"), st.ptr(part_char_func, "part_char_func"), st.ptr(u_z, "u_z"),
        float(U_z_cm_part), n,
        ctypes.c_void_p(partials.data_ptr()), ctypes.c_void_p(out.data_ptr()),
        _lib.stream_handle(),
    )
    v = float(out.item())
    st.finish()
    return v


def compute_force_on_body(
    R, part_char_func, rho_f, brink_lam, u_z, U_z_cm_part, part_vol, dt, diff
):
    """Brinkman penalisation force and unsteady force (compute_forces.py:4-17)."""
    F_pen = rho_f * brink_lam * penalization_force_sum(R, part_char_func, u_z, U_z_cm_part)
    F_un = (diff * part_vol) / dt
    return F_pen, F_un


def max_abs_velocity_sum(u_z, u_r):
    """np.amax(np.fabs(u_z) + np.fabs(u_r)), the CFL reduction of the drivers'
    dt rule (examples/FlowPastSphere/flow_past_sphere.py:150-153)."""
    st = _lib.Staged()
    shape = _lib.same_shape(u_z, u_r)
    n = shape[0] * shape[1]
    partials = st.scratch((_lib.load().axs_reduce_partials_size(n),))
    out = st.scratch((1,))
    _lib.call(
        "axs_max_abs_velocity_sum_f64",
        st.ptr(u_z, "u_z"), st.ptr(u_r, "u_r"), n,
        ctypes.c_void_p(partials.data_ptr()), ctypes.c_void_p(out.data_ptr()),
        _lib.stream_handle(),
    )
    v = float(out.item())
    st.finish()
    return v
