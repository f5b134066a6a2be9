"""Drop-in for the compiled module pyaxisymflow.core.mesh_to_particles
(core/mesh_to_particles_bind.cpp, core/src/mesh_to_particles.hpp:111-131 and
:209-229): M'4 mesh-to-particle interpolation of two fields in 2D, one
gather per particle on the GPU (any particle count and layout)."""
from pyaxisymflow_amd import _lib


def _m2p(input_field_x, input_field_y, particle_positions_x, particle_positions_y,
         output_field_x, output_field_y, delta_x, delta_y, periodic):
    st = _lib.Staged()
    m0, m1 = _lib.same_shape(input_field_x, input_field_y)
    p0, p1 = _lib.same_shape(particle_positions_x, particle_positions_y, output_field_x,
                             output_field_y)
    args = (st.ptr(input_field_x, "input_field_x"), st.ptr(input_field_y, "input_field_y"),
            st.ptr(particle_positions_x, "particle_positions_x"),
            st.ptr(particle_positions_y, "particle_positions_y"),
            st.ptr(output_field_x, "output_field_x", out=True, copy_in=False),
            st.ptr(output_field_y, "output_field_y", out=True, copy_in=False))
    _lib.call("axs_mesh_to_particles_2d_mp4_f64", *args, m0, m1, p0 * p1, float(delta_x),
              float(delta_y), 1 if periodic else 0, _lib.stream_handle())
    st.finish()


def mesh_to_particles_2D_mp4(input_field_x, input_field_y, particle_positions_x,
                             particle_positions_y, output_field_x, output_field_y, delta_x,
                             delta_y):
    """Periodic domain (mesh_to_particles.hpp:111-131)."""
    _m2p(input_field_x, input_field_y, particle_positions_x, particle_positions_y,
         output_field_x, output_field_y, delta_x, delta_y, True)


def mesh_to_particles_2D_unbounded_mp4(input_field_x, input_field_y, particle_positions_x,
                                       particle_positions_y, output_field_x, output_field_y,
                                       delta_x, delta_y):
    """Unbounded domain (mesh_to_particles.hpp:209-229)."""
    _m2p(input_field_x, input_field_y, particle_positions_x, particle_positions_y,
         output_field_x, output_field_y, delta_x, delta_y, False)
