"""Drop-in for pyaxisymflow/kernels/kill_boundary_vorticity_sine.py:4-27."""
from pyaxisymflow_amd import _lib


def _kill(name, vorticity, coord, width, dx):
    st = _lib.Staged()
    nr, nz = _lib.same_shape(vorticity, coord)
    pv = st.ptr(vorticity, "vorticity", out=True)
    pc = st.ptr(coord, "coordinate")
    _lib.call(name, pv, pc, nr, nz, int(width), float(dx), _lib.stream_handle())
    st.finish()


def kill_boundary_vorticity_sine_z(vorticity, Z, width, dx):
    """kill_boundary_vorticity_sine.py:4-14 (sine taper over `width` columns
    at both z ends; the z end is taken as 1, as in the reference)."""
    _kill("axs_kill_boundary_vorticity_sine_z_f64", vorticity, Z, width, dx)


def kill_boundary_vorticity_sine_r(vorticity, R, width, dx):
    """kill_boundary_vorticity_sine.py:17-27 (taper at r_max, taken as 1;
    omega = 0 on the first row)."""
    _kill("axs_kill_boundary_vorticity_sine_r_f64", vorticity, R, width, dx)
