"""GPU parity of the periodic-z drop-ins (the ghost exchange of
kernels/periodic_boundary_ghost_comm.py and the *_periodic stencils) against
the oracle's restatement, through the C ABI.  Bar: bit for bit (copies,
the reference's operation order, no transcendental functions) except the
ENO3 steps, which carry their unbounded forms' bar (bit for bit as well)."""
import numpy as np
import pytest

from oracle import periodic as OP

pytestmark = pytest.mark.gpu

NZ, NR, G = 100, 37, 2


def fields(seed=0, n=3):
    rng = np.random.default_rng(seed)
    dx = 1.0 / NZ
    z = (np.arange(NZ) + 0.5) * dx
    r = (np.arange(NR) + 0.5) * dx
    Z, R = np.meshgrid(z, r)
    return dx, Z, R, [rng.standard_normal((NR, NZ)) for _ in range(n)]


def same(*pairs):
    return all(np.array_equal(a, b) for a, b in pairs)


def test_ghost_exchange_bitexact():
    from pyaxisymflow_amd.kernels import periodic_boundary_ghost_comm as P

    dx, Z, R, (f,) = fields(n=1)
    for gpu_gen, cpu_gen in ((P.gen_periodic_boundary_ghost_comm(G),
                              OP.gen_periodic_boundary_ghost_comm(G)),
                             (P.gen_periodic_boundary_ghost_comm_eta(G, 1.0, dx),
                              OP.gen_periodic_boundary_ghost_comm_eta(G, 1.0, dx))):
        a, b = f.copy(), f.copy()
        gpu_gen(a)
        cpu_gen(b)
        assert np.array_equal(a, b)


def test_velocity_vorticity_rk2_periodic_bitexact():
    from oracle import kernels as K
    from pyaxisymflow_amd.kernels import periodic_boundary_ghost_comm as P
    from pyaxisymflow_amd.kernels.compute_velocity_from_psi import compute_velocity_from_psi_periodic
    from pyaxisymflow_amd.kernels.compute_vorticity_from_velocity import (
        compute_vorticity_from_velocity_periodic,
    )
    from pyaxisymflow_amd.kernels.diffusion_RK2 import diffusion_RK2_periodic

    dx, Z, R, (psi, uz, ur) = fields(1)
    gc, cc = P.gen_periodic_boundary_ghost_comm(G), OP.gen_periodic_boundary_ghost_comm(G)
    p1, p2 = psi.copy(), psi.copy()
    a, b, c, d = (np.zeros_like(psi) for _ in range(4))
    compute_velocity_from_psi_periodic(a, b, p1, R, dx, gc)
    OP.compute_velocity_from_psi_periodic(c, d, p2, R, dx, cc)
    assert same((a, c), (b, d), (p1, p2))
    u1, u2, v1, v2 = uz.copy(), uz.copy(), ur.copy(), ur.copy()
    w1, w2 = np.zeros_like(uz), np.zeros_like(uz)
    compute_vorticity_from_velocity_periodic(w1, u1, v1, dx, gc)
    OP.compute_vorticity_from_velocity_periodic(w2, u2, v2, dx, cc)
    assert same((w1, w2), (u1, u2), (v1, v2))
    x1, x2 = psi.copy(), psi.copy()
    t1, t2 = np.zeros_like(psi), np.zeros_like(psi)
    diffusion_RK2_periodic(x1, t1, R, 0.01, 1e-3, dx, gc)
    OP.diffusion_RK2_periodic(x2, t2, R, 0.01, 1e-3, dx, cc)
    # RK2's R**-2 is numpy pow against 1 / (r r): within an ulp (as the unbounded form)
    assert np.max(np.abs(x1 - x2)) <= 1e-15 * np.max(np.abs(x2))
    with pytest.raises(TypeError):
        diffusion_RK2_periodic(x1, t1, R, 0.01, 1e-3, dx, lambda f: None)


def test_eno3_and_elasto_periodic_bitexact():
    from pyaxisymflow_amd.elasto_kernels.advect_refmap_via_eno3 import (
        gen_advect_refmap_via_eno3_periodic,
    )
    from pyaxisymflow_amd.elasto_kernels.div_tau import update_vorticity_from_solid_stress_periodic
    from pyaxisymflow_amd.elasto_kernels.solid_sigma import solid_sigma_periodic
    from pyaxisymflow_amd.kernels import periodic_boundary_ghost_comm as P
    from pyaxisymflow_amd.kernels.advect_vorticity_via_eno3 import (
        gen_advect_vorticity_via_eno3_periodic,
    )

    dx, Z, R, (w, uz, ur, t11, t12, t22) = fields(2, n=6)
    gc, cc = P.gen_periodic_boundary_ghost_comm(G), OP.gen_periodic_boundary_ghost_comm(G)
    ge = P.gen_periodic_boundary_ghost_comm_eta(G, 1.0, dx)
    ce = OP.gen_periodic_boundary_ghost_comm_eta(G, 1.0, dx)
    dt = 0.2 * dx
    a = [w.copy(), uz.copy(), ur.copy()]
    b = [w.copy(), uz.copy(), ur.copy()]
    gen_advect_vorticity_via_eno3_periodic(dx, NR, NZ, gc)(a[0], a[1], a[2], dt)
    OP.advect_vorticity_via_eno3_periodic(b[0], b[1], b[2], dt, dx, cc)
    assert same(*zip(a, b))
    e1, e2 = Z + 0.01 * w, R + 0.01 * uz
    a = [e1.copy(), e2.copy(), uz.copy(), ur.copy()]
    b = [e1.copy(), e2.copy(), uz.copy(), ur.copy()]
    gen_advect_refmap_via_eno3_periodic(dx, NR, NZ, gc, ge)(*a, dt)
    OP.advect_refmap_via_eno3_periodic(*b, dt, dx, cc, ce)
    assert same(*zip(a, b))
    outs_a = [np.zeros_like(w) for _ in range(7)]
    outs_b = [np.zeros_like(w) for _ in range(7)]
    ea, eb = [e1.copy(), e2.copy()], [e1.copy(), e2.copy()]
    solid_sigma_periodic(*outs_a[:3], 0.7, dx, ea[0], ea[1], *outs_a[3:], gc, ge)
    OP.solid_sigma_periodic(*outs_b[:3], 0.7, dx, eb[0], eb[1], *outs_b[3:], cc, ce)
    assert same(*zip(outs_a, outs_b), *zip(ea, eb))
    a = [w.copy(), np.zeros_like(w), np.zeros_like(w), t11.copy(), t12.copy(), t22.copy()]
    b = [w.copy(), np.zeros_like(w), np.zeros_like(w), t11.copy(), t12.copy(), t22.copy()]
    update_vorticity_from_solid_stress_periodic(*a, R, 1e-3, dx, gc)
    OP.update_vorticity_from_solid_stress_periodic(*b, R, 1e-3, dx, cc)
    assert same(*zip(a, b))
