"""Drop-in for pyaxisymflow/kernels/smooth_Heaviside.py:5-14."""
from pyaxisymflow_amd import _lib


def smooth_Heaviside(H, phi, blend_w):
    """Smooth Heaviside of the level set phi with half-width blend_w
    (smooth_Heaviside.py:5-14).  The comparisons phi >= w and |phi| < w are
    made in FP64 on the same phi, so the mask support is bit-exact."""
    st = _lib.Staged()
    shape = _lib.same_shape(H, phi)
    _lib.call(
        "axs_smooth_heaviside_f64",
        st.ptr(H, "H", out=True, copy_in=False),
        st.ptr(phi, "phi"),
        float(blend_w), shape[0] * shape[1], _lib.stream_handle(),
    )
    st.finish()
