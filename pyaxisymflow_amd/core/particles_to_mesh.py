"""Drop-in for the compiled module pyaxisymflow.core.particles_to_mesh
(core/particles_to_mesh_bind.cpp:35-90, core/src/particles_to_mesh.hpp:71-91
and :165-185): M'4 particle-to-mesh interpolation in 2D.

Arguments are float64 2-D arrays as the reference's pybind11 `noconvert`
binding takes them (numpy arrays are staged to the GPU and back; ROCm torch
tensors are used in place).  The particle arrays must have the mesh's shape
(the remeshing step's particles start on the cell centres): each mesh cell
gathers from the particles within D + 2 of it, D the largest displacement
of a particle's nearest mesh index from its home cell, which must be at most
4 cells (NotImplementedError otherwise).  Results equal the reference's
sequential scatter bit for bit."""
import ctypes

import numpy as np

from pyaxisymflow_amd import _lib


def _p2m(particle_positions_x, particle_positions_y, input_field_at_particle_positions,
         output_field_at_mesh, delta_x, delta_y, periodic):
    st = _lib.Staged()
    shape = _lib.same_shape(particle_positions_x, particle_positions_y,
                            input_field_at_particle_positions, output_field_at_mesh)
    px = st.ptr(particle_positions_x, "particle_positions_x")
    py = st.ptr(particle_positions_y, "particle_positions_y")
    q = st.ptr(input_field_at_particle_positions, "input_field_at_particle_positions")
    out = st.ptr(output_field_at_mesh, "output_field_at_mesh", out=True, copy_in=False)
    disp = ctypes.c_int(0)
    _lib.call("axs_particles_to_mesh_2d_mp4_f64", px, py, q, out, shape[0], shape[1],
              float(delta_x), float(delta_y), 1 if periodic else 0, ctypes.byref(disp),
              _lib.stream_handle())
    st.finish()


def particles_to_mesh_2D_mp4(particle_positions_x, particle_positions_y,
                             input_field_at_particle_positions, output_field_at_mesh, delta_x,
                             delta_y):
    """Periodic domain (particles_to_mesh.hpp:71-91)."""
    _p2m(particle_positions_x, particle_positions_y, input_field_at_particle_positions,
         output_field_at_mesh, delta_x, delta_y, True)


def particles_to_mesh_2D_unbounded_mp4(particle_positions_x, particle_positions_y,
                                       input_field_at_particle_positions, output_field_at_mesh,
                                       delta_x, delta_y):
    """Unbounded domain: stencil cells outside the mesh are dropped
    (particles_to_mesh.hpp:165-185)."""
    _p2m(particle_positions_x, particle_positions_y, input_field_at_particle_positions,
         output_field_at_mesh, delta_x, delta_y, False)
