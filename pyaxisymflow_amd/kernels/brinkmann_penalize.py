"""Drop-in for pyaxisymflow/kernels/brinkmann_penalize.py:4-16."""
from pyaxisymflow_amd import _lib


def brinkmann_penalize(
    lam, dt, char_func, U_z, U_r, grid_u_z, grid_u_r, penalized_u_z, penalized_u_r
):
    """Implicit Brinkman penalisation
    pen_u = (u + lam dt chi U) / (1 + lam dt chi)   (brinkmann_penalize.py:4-16)."""
    st = _lib.Staged()
    shape = _lib.same_shape(char_func, grid_u_z, grid_u_r, penalized_u_z, penalized_u_r)
    n = shape[0] * shape[1]
    _lib.call(
        "axs_brinkmann_penalize_f64",
        float(lam), float(dt),
        st.ptr(char_func, "char_func"),
        float(U_z), float(U_r),
        st.ptr(grid_u_z, "grid_u_z"),
        st.ptr(grid_u_r, "grid_u_r"),
        st.ptr(penalized_u_z, "penalized_u_z", out=True, copy_in=False),
        st.ptr(penalized_u_r, "penalized_u_r", out=True, copy_in=False),
        n, _lib.stream_handle(),
    )
    st.finish()
