// Driver that compiles the reference's header-only particle-mesh
// interpolation (pyaxisymflow/core/src/particles_to_mesh.hpp and
// mesh_to_particles.hpp, read in place under /root/reference) into
// oracle/_ref/libref_particles.so with a C ABI, so the oracle's restatement
// (oracle/particles.py) can be checked against the reference itself.
// TEST INFRASTRUCTURE ONLY.
#include <stdlib.h>

#include <algorithm>

#include "mesh_to_particles.hpp"
#include "particles_to_mesh.hpp"

extern "C" {

// particles_to_mesh.hpp:71-91 (periodic) and :165-185 (unbounded), <double>,
// as bound by core/particles_to_mesh_bind.cpp:35-90: particle arrays (p0, p1),
// mesh (m0, m1), row-major.
void ref_p2m_2d_mp4(const double* px, const double* py, const double* q, double* out, int p0,
                    int p1, int m0, int m1, double dx, double dy, int periodic) {
  if (periodic)
    particles_to_mesh_2D_mp4<double>(px, p0, p1, py, p0, p1, q, p0, p1, out, m0, m1, dx, dy);
  else
    particles_to_mesh_2D_unbounded_mp4<double>(px, p0, p1, py, p0, p1, q, p0, p1, out, m0, m1,
                                                dx, dy);
}

// mesh_to_particles.hpp:111-131 (periodic) and :209-229 (unbounded), <double>.
void ref_m2p_2d_mp4(const double* fx, const double* fy, const double* px, const double* py,
                    double* ox, double* oy, int m0, int m1, int p0, int p1, double dx, double dy,
                    int periodic) {
  if (periodic)
    mesh_to_particles_2D_mp4<double>(fx, m0, m1, fy, m0, m1, px, p0, p1, py, p0, p1, ox, p0, p1,
                                     oy, p0, p1, dx, dy);
  else
    mesh_to_particles_2D_unbounded_mp4<double>(fx, m0, m1, fy, m0, m1, px, p0, p1, py, p0, p1, ox,
                                               p0, p1, oy, p0, p1, dx, dy);
}
}
