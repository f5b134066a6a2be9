"""Writes tests/golden/particles_mp4.npz from the reference's own compiled
headers (oracle/_ref/libref_particles.so, built by `make -C oracle` from
core/src/particles_to_mesh.hpp and mesh_to_particles.hpp in the reference
tree).  Run in the build container:
    python tests/golden/make_particles_golden.py

Cases (seeded): P2M periodic and unbounded of particles displaced from the
cell centres of a 24 x 40 mesh by up to 1.6 cells, some leaving the mesh;
M2P periodic and unbounded of two fields at 300 scattered particles, some
outside the mesh."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle", "_ref",
                   "libref_particles.so")


def main():
    lib = ctypes.CDLL(LIB)
    dp = ctypes.POINTER(ctypes.c_double)

    def ptr(a):
        return a.ctypes.data_as(dp)

    rng = np.random.default_rng(2024)
    m0, m1 = 24, 40
    dx, dy = 1.0 / m1, 1.3 / m1
    x = (np.arange(m1) + 0.5) * dx
    y = (np.arange(m0) + 0.5) * dy
    X, Y = np.meshgrid(x, y)
    px = X + rng.uniform(-1.6, 1.6, X.shape) * dx
    py = Y + rng.uniform(-1.6, 1.6, X.shape) * dy
    q = rng.standard_normal(X.shape)
    out = {}
    for per, name in ((1, "periodic"), (0, "unbounded")):
        o = np.zeros((m0, m1))
        lib.ref_p2m_2d_mp4(ptr(px), ptr(py), ptr(q), ptr(o), m0, m1, m0, m1, ctypes.c_double(dx),
                           ctypes.c_double(dy), per)
        out[f"p2m_{name}"] = o
    fx, fy = rng.standard_normal((m0, m1)), rng.standard_normal((m0, m1))
    qx = rng.uniform(-0.1, 1.1, (15, 20))
    qy = rng.uniform(-0.1, 1.4, (15, 20))
    for per, name in ((1, "periodic"), (0, "unbounded")):
        ox, oy = np.zeros(qx.shape), np.zeros(qx.shape)
        lib.ref_m2p_2d_mp4(ptr(fx), ptr(fy), ptr(qx), ptr(qy), ptr(ox), ptr(oy), m0, m1,
                           qx.shape[0], qx.shape[1], ctypes.c_double(dx), ctypes.c_double(dy),
                           per)
        out[f"m2p_{name}_x"], out[f"m2p_{name}_y"] = ox, oy
    np.savez_compressed(os.path.join(HERE, "particles_mp4.npz"), px=px, py=py, q=q, dx=dx, dy=dy,
                        fx=fx, fy=fy, qx=qx, qy=qy, **out)


if __name__ == "__main__":
    main()
