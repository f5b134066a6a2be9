"""The periodic rigid-slab oracle (oracle/drivers.py PeriodicRigidSlabCase,
periodic_rigid_slab.py:33-158) and its closed-form theory
(womersley_wall_driven, theory_rigid_slab.py:9-23).  CPU only.

* The theory is pinned by what fixes it mathematically: it equals the wall
  velocity U_0 cos(omega t) at r = R_tube, is regular on the axis, and
  satisfies the axisymmetric diffusion equation u_t = nu (u_rr + u_r / r)
  (checked by finite differences with no formula shared with the
  restatement); a planted sign error in the time factor fails that check.
* The driver's two Poisson routes (exact DFT + banded, and the reference's
  dense eig route) agree to 1e-10 over 50 steps; the ghost columns of psi
  hold their periodic sources after a step; dt is the diffusive limit
  0.9 dx^2 / 4 / nu while the flow is slower than the CFL bound.
"""
import numpy as np
import pytest

from oracle.drivers import PeriodicRigidSlabCase, womersley_wall_driven

OMEGA, NU, R_TUBE, U0 = 2 * np.pi * 8.0, 0.02, 0.25, 1.0


def _pde_residual(u):
    r = np.linspace(0.02, 0.24, 221)
    h, dt = r[1] - r[0], 1e-5
    t = 0.37
    f = u(r, t)
    ut = (u(r, t + dt) - u(r, t - dt)) / (2 * dt)
    urr = (f[2:] - 2 * f[1:-1] + f[:-2]) / h**2
    ur = (f[2:] - f[:-2]) / (2 * h)
    res = ut[1:-1] - NU * (urr + ur / r[1:-1])
    return np.max(np.abs(res)) / np.max(np.abs(ut))


def test_theory_wall_value_axis_and_pde():
    u = lambda r, t: womersley_wall_driven(r, t, U0, OMEGA, NU, R_TUBE)  # noqa: E731
    for t in (0.0, 0.013, 0.2, 1.7):
        assert abs(u(np.array([R_TUBE]), t)[0] - U0 * np.cos(OMEGA * t)) < 1e-12
    assert np.isfinite(u(np.array([0.0]), 0.1)).all()
    assert _pde_residual(u) < 1e-3


def test_planted_time_sign_error_fails_pde():
    from scipy.special import jv

    def wrong(r, t):
        lam = np.sqrt(1j * (OMEGA / NU))
        A = 0.5 / jv(0, lam * R_TUBE)
        return U0 * np.real(2 * (A * jv(0, lam * r)) * np.exp(+1j * OMEGA * t))

    assert _pde_residual(wrong) > 0.1


def test_slab_routes_agree_and_ghosts():
    a = PeriodicRigidSlabCase(solver="exact")
    b = PeriodicRigidSlabCase(solver="eig")
    for _ in range(50):
        a.step()
        b.step()
    # the flow is independent of z, so the exact route's u_r is exactly zero:
    # it is measured against u_z's size
    for k, scale in (("vorticity", "vorticity"), ("psi", "psi"), ("u_z", "u_z"), ("u_r", "u_z")):
        x, y = getattr(a, k), getattr(b, k)
        assert np.linalg.norm(x - y) <= 1e-10 * np.linalg.norm(getattr(a, scale)), k
    g = a.g
    assert np.array_equal(a.psi[:, :g], a.psi[:, -2 * g:-g])
    assert np.array_equal(a.psi[:, -g:], a.psi[:, g:2 * g])
    assert a.dts[-1] == pytest.approx(0.9 * a.dx**2 / 4 / a.nu, rel=1e-15)
    assert a.nr == 160 and a.nz == 16
