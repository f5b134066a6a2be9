"""Particle remeshing advection with the M'4 kernel: particle-to-mesh (P2M)
and mesh-to-particle (M2P) interpolation in 2D, periodic and unbounded.
TEST ORACLE ONLY.

Restates
* core/src/particle_kernels/MP4.hpp:14-57 (the two polynomial branches, in
  the header's operation order) and KernelWrapper.hpp:11-92 (stencil points
  -2, -1, 0, +1 take branches 1, 0, 0, 1);
* core/src/interpolation/particles_to_mesh_2D.hpp:13-139 (periodic) and
  :141-326 (unbounded): each particle at X = x / dx sits nearest mesh index
  h = floor(X) + (X >= floor(X) + 0.5); it adds (w_y * w_x) * q to the 4 x 4
  cells h - 2 .. h + 1 (rows outer, columns inner), cell centres at h + s +
  1/2; periodic wraps the indices, unbounded drops cells outside the mesh;
  the output is cleared first and particles are visited in row-major order,
  so each cell's sum runs in particle-index order;
* core/src/interpolation/mesh_to_particles_2D.hpp:12-196 (periodic) and
  :198-427 (unbounded): per particle, partial[sy] = sum_sx w_x * f (columns
  in order) for two fields at once, then sum_sy w_y * partial[sy];
* kernels/advect_particle.py:5-68: push the particles of the doubled
  (2 nr, nz) grid by the mirrored velocity, carry the mirrored vorticity
  (odd about the axis), remesh with P2M, reset the particles to the grid,
  keep the upper half.

The compiled reference headers (oracle/_ref/libref_particles.so, built by
oracle/Makefile) and the fixtures written from them (tests/golden/) pin this
restatement bit for bit.
"""
import numpy as np


def mp4_f0(t):
    """MP4.hpp:26-28, |x| <= 1."""
    return 1.0 + t * t * (-2.5 + 1.5 * t)


def mp4_f1(t):
    """MP4.hpp:35-38, 1 < |x| <= 2."""
    return 2.0 + t * (-4.0 + t * (2.5 - 0.5 * t))


OFFSETS = (-2, -1, 0, 1)
BRANCH = (mp4_f1, mp4_f0, mp4_f0, mp4_f1)


def stencil(pos, d):
    """Nearest mesh index h and the four weights of each particle."""
    X = pos / d
    least = np.floor(X)
    h = (X >= least + 0.5).astype(np.int64) + least.astype(np.int64)
    w = [f(np.fabs(X - ((h + 0.5) + s))) for s, f in zip(OFFSETS, BRANCH)]
    return h, w


def particles_to_mesh_2D(px, py, q, out, dx, dy, periodic):
    """P2M of q at particles (px, py) onto out (ny, nx), in place."""
    ny, nx = out.shape
    hx, wx = stencil(px.ravel(), dx)
    hy, wy = stencil(py.ravel(), dy)
    qv = q.ravel()
    idx = np.empty((qv.size, 16), dtype=np.int64)
    val = np.empty((qv.size, 16))
    ok = np.empty((qv.size, 16), dtype=bool)
    k = 0
    for a, sy in enumerate(OFFSETS):
        for b, sx in enumerate(OFFSETS):
            row, col = hy + sy, hx + sx
            if periodic:
                row = (row + ny) % ny
                col = (col + nx) % nx
                ok[:, k] = True
            else:
                ok[:, k] = (row >= 0) & (row < ny) & (col >= 0) & (col < nx)
            idx[:, k] = np.where(ok[:, k], row * nx + col, 0)
            val[:, k] = (wy[a] * wx[b]) * qv
            k += 1
    flat = np.zeros(ny * nx)
    sel = ok.ravel()
    # np.add.at applies the additions one by one in index-array order:
    # particle-major, then the 16 stencil cells, as the header's loops
    np.add.at(flat, idx.ravel()[sel], val.ravel()[sel])
    out[...] = flat.reshape(ny, nx)


def particles_to_mesh_2D_mp4(px, py, q, out, dx, dy):
    particles_to_mesh_2D(px, py, q, out, dx, dy, periodic=True)


def particles_to_mesh_2D_unbounded_mp4(px, py, q, out, dx, dy):
    particles_to_mesh_2D(px, py, q, out, dx, dy, periodic=False)


def mesh_to_particles_2D(fx, fy, px, py, outx, outy, dx, dy, periodic):
    """M2P of the two mesh fields (fx, fy) to the particles, in place."""
    ny, nx = fx.shape
    hx, wx = stencil(px.ravel(), dx)
    hy, wy = stencil(py.ravel(), dy)
    n = hx.size
    accx = np.zeros(n)
    accy = np.zeros(n)
    fxf, fyf = fx.ravel(), fy.ravel()
    partial = []
    for a, sy in enumerate(OFFSETS):
        sx_sum = np.zeros(n)
        sy_sum = np.zeros(n)
        for b, sx in enumerate(OFFSETS):
            row, col = hy + sy, hx + sx
            if periodic:
                row = (row + ny) % ny
                col = (col + nx) % nx
                ok = np.ones(n, dtype=bool)
            else:
                ok = (row >= 0) & (row < ny) & (col >= 0) & (col < nx)
            m = np.where(ok, row * nx + col, 0)
            sx_sum = np.where(ok, sx_sum + wx[b] * fxf[m], sx_sum)
            sy_sum = np.where(ok, sy_sum + wx[b] * fyf[m], sy_sum)
        partial.append((sx_sum, sy_sum))
    for a in range(4):
        accx = accx + wy[a] * partial[a][0]
        accy = accy + wy[a] * partial[a][1]
    outx[...] = accx.reshape(outx.shape)
    outy[...] = accy.reshape(outy.shape)


def mesh_to_particles_2D_mp4(fx, fy, px, py, outx, outy, dx, dy):
    mesh_to_particles_2D(fx, fy, px, py, outx, outy, dx, dy, periodic=True)


def mesh_to_particles_2D_unbounded_mp4(fx, fy, px, py, outx, outy, dx, dy):
    mesh_to_particles_2D(fx, fy, px, py, outx, outy, dx, dy, periodic=False)


def advect_vorticity_via_particles(z_particles, r_particles, vort_particles, vorticity, Z_double,
                                   R_double, grid_size_r, u_z, u_r, dx, dt, periodic=False):
    """kernels/advect_particle.py:5-35 (unbounded) and :38-68 (periodic)."""
    nr = grid_size_r
    z_particles[nr:, :] += u_z * dt
    z_particles[:nr, :] += np.flip(u_z, axis=0) * dt
    r_particles[nr:, :] += u_r * dt
    r_particles[:nr, :] += -np.flip(u_r, axis=0) * dt
    vort_particles[nr:, :] = vorticity
    vort_particles[:nr, :] = -np.flip(vorticity, axis=0)
    vort_double = 0 * Z_double
    particles_to_mesh_2D(z_particles, r_particles, vort_particles, vort_double, dx, dx, periodic)
    z_particles[...] = Z_double
    r_particles[...] = R_double
    vorticity[...] = vort_double[nr:, :]
