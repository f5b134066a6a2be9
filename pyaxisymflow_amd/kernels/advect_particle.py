"""Drop-in for pyaxisymflow/kernels/advect_particle.py:5-68: vorticity
advected by particles on the doubled (2 nr, nz) grid and remeshed with the
M'4 kernel, every stage on the GPU (axs_advect_vorticity_via_particles_f64).
Same arguments and in-place results as the reference: the particles end
back on (Z_double, R_double), vort_particles holds the mirrored vorticity
the remeshing carried, vorticity the remeshed upper half."""
import ctypes

from pyaxisymflow_amd import _lib


def _advect(z_particles, r_particles, vort_particles, vorticity, Z_double, R_double,
            grid_size_r, u_z, u_r, dx, dt, periodic):
    st = _lib.Staged()
    nr, nz = _lib.same_shape(vorticity, u_z, u_r)
    d = _lib.same_shape(z_particles, r_particles, vort_particles, Z_double, R_double)
    if nr != grid_size_r or d != (2 * nr, nz):
        raise ValueError(f"particle arrays {d} must be (2 grid_size_r, nz) = {(2 * nr, nz)}")
    zp = st.ptr(z_particles, "z_particles", out=True)
    rp = st.ptr(r_particles, "r_particles", out=True)
    wp = st.ptr(vort_particles, "vort_particles", out=True)
    w = st.ptr(vorticity, "vorticity", out=True)
    work = st.scratch(d)
    _lib.call("axs_advect_vorticity_via_particles_f64", zp, rp, wp, w,
              st.ptr(Z_double, "Z_double"), st.ptr(R_double, "R_double"),
              st.ptr(u_z, "u_z"), st.ptr(u_r, "u_r"), ctypes.c_void_p(work.data_ptr()),
              nr, nz, float(dx), float(dt), 1 if periodic else 0, _lib.stream_handle())
    st.finish()


def advect_vorticity_via_particles(z_particles, r_particles, vort_particles, vorticity, Z_double,
                                   R_double, grid_size_r, u_z, u_r, dx, dt):
    """advect_particle.py:5-35 (unbounded remeshing)."""
    _advect(z_particles, r_particles, vort_particles, vorticity, Z_double, R_double, grid_size_r,
            u_z, u_r, dx, dt, False)


def advect_vorticity_via_particles_periodic(z_particles, r_particles, vort_particles, vorticity,
                                            Z_double, R_double, grid_size_r, u_z, u_r, dx, dt):
    """advect_particle.py:38-68 (periodic remeshing)."""
    _advect(z_particles, r_particles, vort_particles, vorticity, Z_double, R_double, grid_size_r,
            u_z, u_r, dx, dt, True)
