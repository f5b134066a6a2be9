"""GPU parity of every drop-in kernel against the CPU oracle, through the C
ABI (pyaxisymflow_amd -> libaxisym_hip.so).

Bars: bit-exact where the reference's arithmetic has no transcendental
functions (the HIP build keeps the reference's operation order and turns
FMA contraction off); otherwise relative error <= 1e-13 (sin, pow) and the
Poisson solve within 1e-10 relative L2 (FP64 target), reported against both
the eig-route oracle and the exact discrete solution."""
import numpy as np
import pytest

from oracle import kernels as K
from oracle import lsq as olsq
from oracle.drivers import gaussian_vorticity, grid, sphere_union_phi
from oracle.stokes import FastDiagonalisationStokesSolver as OracleSolver
from oracle.stokes import exact_stokes_solve, stokes_residual

pytestmark = pytest.mark.gpu

SIZES = [(64, 32), (256, 128), (512, 256)]  # (nz, nr)


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def fields(nz, nr, seed=0, z_len=2.0):
    dx, z, r, Z, R = grid(nz, nr, z_len)
    w = gaussian_vorticity(Z, R, 16, seed, 0.02, 0.1, z_len, nr * dx)
    rng = np.random.default_rng(seed + 100)
    uz = 1.0 + 0.3 * rng.standard_normal((nr, nz))
    ur = 0.3 * rng.standard_normal((nr, nz))
    return dx, z, r, Z, R, w, uz, ur


@pytest.mark.parametrize("nz,nr", SIZES)
def test_brinkmann_bitexact(nz, nr):
    from pyaxisymflow_amd.kernels.brinkmann_penalize import brinkmann_penalize

    dx, z, r, Z, R, w, uz, ur = fields(nz, nr)
    phi = sphere_union_phi(Z, R, 8, 1, 0.02, 0.06, 0.2, 1.8)
    chi = np.zeros_like(phi)
    K.smooth_Heaviside(chi, phi, np.sqrt(2) * dx)
    a, b = np.zeros_like(uz), np.zeros_like(uz)
    c, d = np.zeros_like(uz), np.zeros_like(uz)
    brinkmann_penalize(1e4, 1.3e-4, chi, 0.7, -0.2, uz, ur, a, b)
    K.brinkmann_penalize(1e4, 1.3e-4, chi, 0.7, -0.2, uz, ur, c, d)
    assert np.array_equal(a, c) and np.array_equal(b, d)


def test_brinkmann_reference_kat():
    """tests/test_kernels/test_brinkmann_penalize.py:6-26 on the GPU."""
    from pyaxisymflow_amd.kernels.brinkmann_penalize import brinkmann_penalize

    n = 16
    chi = 4 * np.ones((n, n))
    z0 = np.zeros((n, n))
    pz, pr = np.zeros((n, n)), np.zeros((n, n))
    brinkmann_penalize(2.0, 3.0, chi, 1.0, 2.0, z0, z0.copy(), pz, pr)
    np.testing.assert_allclose(pz, 2 * 3 * 1 * chi / (1 + 2 * 3 * chi))
    np.testing.assert_allclose(pr, 2 * 3 * 2 * chi / (1 + 2 * 3 * chi))


@pytest.mark.parametrize("nz,nr", SIZES)
def test_heaviside(nz, nr):
    from pyaxisymflow_amd.kernels.smooth_Heaviside import smooth_Heaviside

    dx, z, r, Z, R, *_ = fields(nz, nr)
    phi = sphere_union_phi(Z, R, 8, 2, 0.02, 0.06, 0.2, 1.8)
    w = np.sqrt(2) * dx
    h, ho = np.empty_like(phi), np.empty_like(phi)
    smooth_Heaviside(h, phi, w)
    K.smooth_Heaviside(ho, phi, w)
    # mask support (the comparisons) bit-exact; blend values to round-off
    assert np.array_equal(h == 0, ho == 0) and np.array_equal(h == 1, ho == 1)
    assert np.array_equal(h > 0.5, ho > 0.5)
    np.testing.assert_allclose(h, ho, rtol=0, atol=2e-16)


@pytest.mark.parametrize("nz,nr", SIZES)
def test_kill_taper(nz, nr):
    from pyaxisymflow_amd.kernels.kill_boundary_vorticity_sine import (
        kill_boundary_vorticity_sine_r,
        kill_boundary_vorticity_sine_z,
    )

    dx, z, r, Z, R, w, *_ = fields(nz, nr)
    a, b = w.copy(), w.copy()
    kill_boundary_vorticity_sine_z(a, Z, 3, dx)
    kill_boundary_vorticity_sine_r(a, R, 3, dx)
    K.kill_boundary_vorticity_sine_z(b, Z, 3, dx)
    K.kill_boundary_vorticity_sine_r(b, R, 3, dx)
    np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-300)
    assert np.array_equal(a[:, 3:-3][1:-3], w[:, 3:-3][1:-3])


@pytest.mark.parametrize("nz,nr", SIZES)
def test_velocity_vorticity_bitexact(nz, nr):
    from pyaxisymflow_amd.kernels.compute_velocity_from_psi import compute_velocity_from_psi_unb
    from pyaxisymflow_amd.kernels.compute_vorticity_from_velocity import (
        compute_vorticity_from_velocity_unb,
    )

    dx, z, r, Z, R, w, uz, ur = fields(nz, nr)
    psi = np.sin(3 * Z) * R**2 + 0.1 * w
    a, b, c, d = (np.zeros_like(psi) for _ in range(4))
    compute_velocity_from_psi_unb(a, b, psi, R, dx)
    K.compute_velocity_from_psi_unb(c, d, psi, R, dx)
    assert np.array_equal(a, c) and np.array_equal(b, d)
    v1 = np.full_like(psi, 3.0)
    v2 = v1.copy()
    compute_vorticity_from_velocity_unb(v1, uz, ur, dx)
    K.compute_vorticity_from_velocity_unb(v2, uz, ur, dx)
    assert np.array_equal(v1, v2)


@pytest.mark.parametrize("nz,nr", SIZES)
def test_diffusion_rk2(nz, nr):
    from pyaxisymflow_amd.kernels.diffusion_RK2 import diffusion_RK2_unb

    dx, z, r, Z, R, w, *_ = fields(nz, nr)
    a, b = w.copy(), w.copy()
    ta, tb = np.zeros_like(w), np.zeros_like(w)
    nu, dt = 1.5e-3, 0.9 * dx**2 / 4 / 1.5e-3
    diffusion_RK2_unb(a, ta, R, nu, dt, dx)
    K.diffusion_RK2_unb(b, tb, R, nu, dt, dx)
    # only r**-2 (pow vs 1/(r*r)) may differ, by an ulp
    assert rel_l2(ta, tb) < 1e-15 and rel_l2(a, b) < 1e-15
    np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-13 * np.abs(b).max())


@pytest.mark.parametrize("nz,nr", SIZES)
def test_eno3_vorticity_bitexact(nz, nr):
    from pyaxisymflow_amd.kernels.advect_vorticity_via_eno3 import gen_advect_vorticity_via_eno3

    dx, z, r, Z, R, w, uz, ur = fields(nz, nr)
    adv = gen_advect_vorticity_via_eno3(dx, nr, nz)
    a, b = w.copy(), w.copy()
    dt = 0.1 * dx / np.abs(uz + ur).max()
    adv(a, uz, ur, dt)
    K.advect_vorticity_via_eno3(b, uz, ur, dt, dx)
    assert np.array_equal(a, b)
    adv(a, uz, ur, dt)  # the closure is reusable
    K.advect_vorticity_via_eno3(b, uz, ur, dt, dx)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("nz,nr", SIZES)
def test_eno3_refmap_bitexact(nz, nr):
    from pyaxisymflow_amd.elasto_kernels.advect_refmap_via_eno3 import gen_advect_refmap_via_eno3

    dx, z, r, Z, R, w, uz, ur = fields(nz, nr)
    adv = gen_advect_refmap_via_eno3(dx, nr, nz)
    e1, e2 = Z + 0.01 * w, R - 0.01 * w
    a1, a2, b1, b2 = e1.copy(), e2.copy(), e1.copy(), e2.copy()
    dt = 0.1 * dx / np.abs(uz + ur).max()
    adv(a1, a2, uz, ur, dt)
    K.advect_refmap_via_eno3(b1, b2, uz, ur, dt, dx)
    assert np.array_equal(a1, b1) and np.array_equal(a2, b2)


@pytest.mark.parametrize("nz,nr", SIZES)
def test_solid_sigma_and_div_tau_bitexact(nz, nr):
    from pyaxisymflow_amd.elasto_kernels.div_tau import update_vorticity_from_solid_stress
    from pyaxisymflow_amd.elasto_kernels.solid_sigma import solid_sigma

    dx, z, r, Z, R, w, uz, ur = fields(nz, nr)
    e1, e2 = Z + 0.01 * w, R - 0.02 * np.cos(Z) * R
    rng = np.random.default_rng(9)
    # pre-filled scratch: never-written entries must carry through
    g_gpu = [rng.standard_normal(Z.shape) for _ in range(4)]
    g_cpu = [g.copy() for g in g_gpu]
    s_gpu = [np.zeros_like(Z) for _ in range(3)]
    s_cpu = [np.zeros_like(Z) for _ in range(3)]
    solid_sigma(*s_gpu, 1.7, dx, e1, e2, *g_gpu)
    K.solid_sigma(*s_cpu, 1.7, dx, e1, e2, *g_cpu)
    for x, y in zip(s_gpu + g_gpu, s_cpu + g_cpu):
        assert np.array_equal(x, y)
    wa, wb = w.copy(), w.copy()
    ta = [rng.standard_normal(Z.shape) for _ in range(2)]
    tb = [t.copy() for t in ta]
    update_vorticity_from_solid_stress(wa, ta[0], ta[1], *s_gpu, R, 1e-3, dx)
    K.update_vorticity_from_solid_stress(wb, tb[0], tb[1], *s_cpu, R, 1e-3, dx)
    assert np.array_equal(wa, wb)
    assert np.array_equal(ta[0], tb[0]) and np.array_equal(ta[1], tb[1])


@pytest.mark.parametrize("nz,nr", SIZES)
def test_reductions(nz, nr):
    from pyaxisymflow_amd.kernels.compute_forces import compute_force_on_body, max_abs_velocity_sum

    dx, z, r, Z, R, w, uz, ur = fields(nz, nr)
    phi = sphere_union_phi(Z, R, 8, 1, 0.02, 0.06, 0.2, 1.8)
    chi = np.zeros_like(phi)
    K.smooth_Heaviside(chi, phi, np.sqrt(2) * dx)
    g = compute_force_on_body(R, chi, 1.0, 1e4, uz, 0.3, 0.01, 1e-4, 0.2)
    c = K.compute_force_on_body(R, chi, 1.0, 1e4, uz, 0.3, 0.01, 1e-4, 0.2)
    np.testing.assert_allclose(g[0], c[0], rtol=1e-13)
    assert g[1] == c[1]
    assert max_abs_velocity_sum(uz, ur) == K.max_abs_velocity_sum(uz, ur)


@pytest.mark.parametrize("nz,nr", SIZES)
def test_stokes_solve(nz, nr):
    """Relative L2 of psi <= 1e-10 against the eig-route oracle and <= 1e-12
    against the exact discrete solution; the residual of the discrete system
    is no worse than twice the CPU exact route's and below the eig route's."""
    from pyaxisymflow_amd.kernels.FastDiagonalisationStokesSolver import (
        FastDiagonalisationStokesSolver,
    )

    dx = 1.0 / nz
    z = np.linspace(dx / 2, 1 - dx / 2, nz)
    r = np.linspace(dx / 2, nr * dx - dx / 2, nr)
    Z, R = np.meshgrid(z, r)
    w = gaussian_vorticity(Z, R, 16, 3, 0.02, 0.1, 1.0, nr * dx)
    gpu = FastDiagonalisationStokesSolver(nr, nz, dx)
    psi = np.zeros_like(w)
    gpu.solve(solution_field=psi, rhs_field=w)
    exact = exact_stokes_solve(w, dx)
    assert rel_l2(psi, exact) < 1e-12
    eig = np.zeros_like(w)
    OracleSolver(nr, nz, dx).solve(eig, w)
    assert rel_l2(psi, eig) < 1e-10
    res = stokes_residual(psi, w, dx)
    assert res <= max(2 * stokes_residual(exact, w, dx), 1e-13)
    assert res < stokes_residual(eig, w, dx)


@pytest.mark.parametrize("nz,nr", [(96, 48), (384, 192), (768, 384), (160, 80)])
def test_stokes_mixed_radix(nz, nr):
    """nz/2 with factors 3 and 5 (the batched workload's 384 and 768 grids)."""
    from pyaxisymflow_amd.kernels.FastDiagonalisationStokesSolver import (
        FastDiagonalisationStokesSolver,
    )

    dx = 1.0 / nz
    w = np.random.default_rng(nz).standard_normal((nr, nz))
    psi = np.zeros_like(w)
    FastDiagonalisationStokesSolver(nr, nz, dx).solve(psi, w)
    assert rel_l2(psi, exact_stokes_solve(w, dx)) < 1e-12


@pytest.mark.parametrize("nz,nr", [(256, 40), (384, 30), (512, 50), (768, 40), (1024, 64),
                                   (1024, 45), (2048, 64), (2048, 131), (4096, 64), (4096, 37)])
def test_stokes_register_fft_rows(nz, nr):
    """nz = 256 to 2048 run the one-wave-per-row register FFTs
    (axs_fft1024.hpp: R1 x 8 x 8 points with R1 = 2, 3, 4, 6, 8, and 16 x 64),
    nz = 4096 two waves per row (one 1024-point FFT each and a radix-2
    join): against the exact discrete solution and its residual."""
    from pyaxisymflow_amd.kernels.FastDiagonalisationStokesSolver import (
        FastDiagonalisationStokesSolver,
    )

    dx = 1.0 / nz
    z = np.linspace(dx / 2, 1 - dx / 2, nz)
    r = np.linspace(dx / 2, nr * dx - dx / 2, nr)
    Z, R = np.meshgrid(z, r)
    w = gaussian_vorticity(Z, R, 16, 5, 0.005, 0.02, 1.0, nr * dx)
    w += 0.1 * np.random.default_rng(nr).standard_normal(w.shape)
    psi = np.zeros_like(w)
    FastDiagonalisationStokesSolver(nr, nz, dx).solve(solution_field=psi, rhs_field=w)
    exact = exact_stokes_solve(w, dx)
    print(f"nz={nz} nr={nr}: rel L2 vs exact {rel_l2(psi, exact):.2e}", flush=True)
    assert rel_l2(psi, exact) < 1e-12
    assert stokes_residual(psi, w, dx) <= max(2 * stokes_residual(exact, w, dx), 1e-13)


def test_stokes_rejects_unsupported():
    from pyaxisymflow_amd.kernels.FastDiagonalisationStokesSolver import (
        FastDiagonalisationStokesSolver,
    )

    with pytest.raises(NotImplementedError):
        # nz / 2 = 4999 is prime and nz > 8192: beyond the row transforms' LDS
        FastDiagonalisationStokesSolver(8, 9998, 0.1)
    # the three bc_types of FastDiagonalisationStokesSolver.py:78-97 are
    # supported (periodic z: test_gpu_periodic.py); any other is refused
    with pytest.raises(NotImplementedError):
        FastDiagonalisationStokesSolver(8, 16, 0.1, bc_type="homogenous_dirichlet_along_z_and_r")


@pytest.mark.parametrize("nz,seed", [(64, 1), (128, 2), (256, 3)])
def test_lsq_extrapolation_bitexact(nz, seed):
    from pyaxisymflow_amd.core.extrapolate_using_least_squares import (
        extrapolate_using_least_squares_till_first_order,
    )
    from test_oracle_lsq import _soft_sphere

    cf, tf, e1, e2, z = _soft_sphere(nz, seed)
    a = [cf.copy(), tf.copy(), e1.copy(), e2.copy()]
    b = [cf.copy(), tf.copy(), e1.copy(), e2.copy()]
    ng = extrapolate_using_least_squares_till_first_order(*a, z, z)
    nc = olsq.extrapolate_using_least_squares_till_first_order(*b, z, z)
    assert ng == nc
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_lsq_golden_fixture():
    import os

    from pyaxisymflow_amd.core.extrapolate_using_least_squares import (
        extrapolate_using_least_squares_till_first_order,
    )

    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "lsq_soft_sphere_nz64.npz"))
    cf, tf, ex, ey = (g[k].copy() for k in ("cur_in", "tgt", "ex_in", "ey_in"))
    extrapolate_using_least_squares_till_first_order(cf, tf, ex, ey, g["gx"], g["gy"])
    assert np.array_equal(cf, g["cur_out"])
    assert np.array_equal(ex, g["ex_out"]) and np.array_equal(ey, g["ey_out"])


def test_torch_tensors_in_place():
    """Device tensors are used in place on the current stream (no copies)."""
    import torch

    from pyaxisymflow_amd.kernels.compute_velocity_from_psi import compute_velocity_from_psi_unb

    nz, nr = 128, 64
    dx, z, r, Z, R, w, uz, ur = fields(nz, nr)
    psi = np.cos(2 * Z) * R**2
    t = lambda a: torch.from_numpy(a).cuda()
    tuz, tur = torch.zeros_like(t(psi)), torch.zeros_like(t(psi))
    compute_velocity_from_psi_unb(tuz, tur, t(psi), t(R), dx)
    c, d = np.zeros_like(psi), np.zeros_like(psi)
    K.compute_velocity_from_psi_unb(c, d, psi, R, dx)
    assert np.array_equal(tuz.cpu().numpy(), c) and np.array_equal(tur.cpu().numpy(), d)
