"""Grids whose nz the half-length transforms do not cover: odd nz, or nz / 2
with a prime factor above 5 (7, 11, 97, 1001 = 7 x 11 x 13, 2047 = 23 x 89).
The reference's dense eig route takes any (nr, nz)
(kernels/FastDiagonalisationStokesSolver.py:41-128,
kernels/implicit_diffusion_solver.py:74-106,
kernels/FastDiagonalisationPotentialSolver.py:41-120); here the row passes
run the any-length transforms (Stockham for 2,3,5-smooth lengths, Bluestein's
chirp-z form otherwise).

Bars (FP64): psi within 1e-12 relative L2 of the exact discrete solution
(DCT / DFT / DST-I by scipy.fft plus banded radial solves) and 1e-10 of the
eig route where that route is run; the steppers within the FP64 bar of the
oracle's driver loops."""
import numpy as np
import pytest

from oracle.drivers import gaussian_vorticity
from oracle.stokes import FastDiagonalisationStokesSolver as OracleSolver
from oracle.stokes import exact_stokes_solve, stokes_residual

pytestmark = pytest.mark.gpu

NEUMANN = "homogenous_neumann_along_z_and_r"
PERIODIC = ("homogenous_neumann_along_r_and_periodic_along_z",
            "homogenous_dirichlet_along_r_and_periodic_along_z")


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _rhs(nz, nr, seed=3):
    dx = 1.0 / nz
    z = np.linspace(dx / 2, 1 - dx / 2, nz)
    r = np.linspace(dx / 2, nr * dx - dx / 2, nr)
    Z, R = np.meshgrid(z, r)
    w = gaussian_vorticity(Z, R, 16, seed, 0.02, 0.1, 1.0, nr * dx)
    w += 0.1 * np.random.default_rng(nz + nr).standard_normal(w.shape)
    return dx, w


@pytest.mark.parametrize("nz,nr", [(14, 7), (154, 77), (194, 97), (15, 9), (97, 40),
                                   (2002, 64), (4094, 33), (8192, 16)])
def test_stokes_any_nz(nz, nr):
    from pyaxisymflow_amd.kernels.FastDiagonalisationStokesSolver import (
        FastDiagonalisationStokesSolver,
    )

    dx, w = _rhs(nz, nr)
    print(f"stokes nz={nz} nr={nr}", flush=True)
    psi = np.zeros_like(w)
    FastDiagonalisationStokesSolver(nr, nz, dx).solve(solution_field=psi, rhs_field=w)
    exact = exact_stokes_solve(w, dx)
    e = rel_l2(psi, exact)
    msg = f"stokes {nz}x{nr}: vs exact {e:.2e}"
    assert e < 1e-12, msg
    assert stokes_residual(psi, w, dx) <= max(2 * stokes_residual(exact, w, dx), 1e-13)
    if nz <= 200:
        eig = np.zeros_like(w)
        OracleSolver(nr, nz, dx).solve(eig, w)
        msg += f", vs eig route {rel_l2(psi, eig):.2e}"
        assert rel_l2(psi, eig) < 1e-10, msg
    print(msg, flush=True)


@pytest.mark.parametrize("bc", PERIODIC)
@pytest.mark.parametrize("nz,nr", [(15, 8), (97, 33), (154, 40), (2002, 20)])
def test_periodic_stokes_any_nz(bc, nz, nr):
    from pyaxisymflow_amd.kernels.FastDiagonalisationStokesSolver import (
        FastDiagonalisationStokesSolver,
    )

    dx = 1.0 / nz
    w = np.random.default_rng(nz + nr).standard_normal((nr, nz))
    psi = np.zeros_like(w)
    FastDiagonalisationStokesSolver(nr, nz, dx, bc_type=bc).solve(solution_field=psi, rhs_field=w)
    exact = exact_stokes_solve(w, dx, bc)
    e = rel_l2(psi, exact)
    print(f"periodic {bc[:22]} {nz}x{nr}: vs exact {e:.2e}", flush=True)
    assert e < 1e-12
    assert stokes_residual(psi, w, dx, bc) < 1e-12


@pytest.mark.parametrize("nz,nr", [(14, 9), (154, 77), (255, 100), (256, 128), (2048, 64)])
def test_implicit_diffusion_fast_sine_rows(nz, nr):
    """The implicit step's rows are now sine transforms through a DFT of
    nz + 1 points (Bluestein for 257 and 2049, Stockham for 15 and 256):
    against the exact discrete solution and the stencil residual."""
    from oracle import implicit_diffusion as O
    from pyaxisymflow_amd.kernels.implicit_diffusion_solver import ImplicitEulerDiffusionStepper

    dx = 1.0 / nz
    nu = 1.5e-3
    dt = dx**2 / 4 / nu
    w = np.random.default_rng(nz).standard_normal((nr, nz))
    gpu = w.copy()
    ImplicitEulerDiffusionStepper(time_step=dt, kinematic_viscosity=nu, grid_size_r=nr,
                                  grid_size_z=nz, dx=dx).step(vorticity_field=gpu, dt=dt)
    exact = O.exact_implicit_diffusion(w, nu * dt, dx)
    e = rel_l2(gpu, exact)
    print(f"implicit {nz}x{nr}: vs exact {e:.2e}", flush=True)
    assert e < 1e-12
    assert O.implicit_diffusion_residual(gpu, w, nu * dt, dx) < 1e-12


@pytest.mark.parametrize("nz,nr", [(154, 77), (97, 50)])
def test_potential_any_nz(nz, nr):
    """A compatible rhs: phi against the exact route in the same gauge
    (phi_hat_0[nr - 1] = 0) and the stencil residual."""
    from oracle import potential as O
    from pyaxisymflow_amd.kernels.FastDiagonalisationPotentialSolver import (
        FastDiagonalisationPotentialSolver,
    )

    dx = 1.0 / nz
    rhs = O.make_compatible(np.random.default_rng(nz).standard_normal((nr, nz)), dx)
    phi = np.zeros_like(rhs)
    FastDiagonalisationPotentialSolver(nr, nz, dx).solve(solution_field=phi, rhs_field=rhs)
    e = rel_l2(phi, O.exact_potential_solve(rhs, dx))
    res = O.potential_residual(phi, rhs, dx)
    print(f"potential {nz}x{nr}: vs exact {e:.2e}, residual {res:.2e}", flush=True)
    assert e < 1e-12 and res < 1e-12


@pytest.mark.parametrize("nz,nr", [(154, 77), (135, 70)])
def test_rigid_stepper_any_nz(nz, nr):
    """Flow past a sphere on z in [0, 1] with nz = 154 (nz / 2 = 7 x 11) and
    odd nz = 135: 40 steps against the oracle's driver loop."""
    from oracle.drivers import RigidBodyCase, grid
    from oracle.kernels import smooth_Heaviside
    from pyaxisymflow_amd.stepper import RigidBodyStepper

    dx, z, r, Z, R = grid(nz, nr, 1.0)
    chi = np.zeros_like(Z)
    smooth_Heaviside(chi, 0.1 - np.sqrt((Z - 0.4) ** 2 + R**2), dx * 2**0.5)
    w0 = gaussian_vorticity(Z, R, 16, 7, 0.02, 0.1, 1.0, nr * dx)
    kw = dict(U_0=1.0, T_ramp=0.5, r_sph=0.1, vorticity=w0)
    cpu = RigidBodyCase(nz, nr, 1.0, 2e-3, 1e4, chi, solver="exact", **kw)
    gpu = RigidBodyStepper(nz, nr, 1.0, 2e-3, 1e4, chi, **kw)
    for _ in range(40):
        cpu.step()
    gpu.step(40)
    ew, ep = rel_l2(gpu.vorticity(), cpu.vorticity), rel_l2(gpu.psi(), cpu.psi)
    print(f"rigid stepper {nz}x{nr}: omega {ew:.2e} psi {ep:.2e}", flush=True)
    assert ew < 1e-10 and ep < 1e-10
    np.testing.assert_allclose(gpu.drag_history(), np.array(cpu.drag), rtol=1e-10, atol=1e-12)


def test_soft_stepper_nz_not_smooth():
    """The soft sphere on nz = 196 (nz / 2 = 98 = 2 x 7^2), one step against
    the oracle within 1e-12 on every field."""
    from oracle.drivers import SoftSphereCase
    from pyaxisymflow_amd.stepper import SoftSphereStepper

    cpu = SoftSphereCase(grid_size_z=196, solver="exact")
    gpu = SoftSphereStepper(grid_size_z=196)
    cpu.step()
    gpu.step(1)
    e1, e2 = gpu.eta()
    errs = dict(w=rel_l2(gpu.vorticity(), cpu.vorticity), psi=rel_l2(gpu.psi(), cpu.psi),
                e1=rel_l2(e1, cpu.eta1), e2=rel_l2(e2, cpu.eta2))
    print("soft nz=196: " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v < 1e-12, k
