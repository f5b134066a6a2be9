"""The periodic soft-slab oracle (oracle/drivers.py PeriodicSoftSlabCase,
periodic_soft_slab.py:43-404), its static PDE extrapolation
(oracle/pde_extrapolation.py, bounded_static_PDE_extrapolation.py:5-235) and
its closed-form theory (soft_slab_theory, theory_soft_slab.py:15-87).  CPU only.

* The theory is pinned by what fixes it: the wall velocity V cos(omega t) at
  y = L_s + L_f; velocity and shear stress continuous at the interface
  (mu_f dv_f/dy = G du_s/dy + mu_s dv_s/dy, derivatives by finite
  differences); the fluid satisfying v_t = nu_f (v_yy + v_y / y) and the
  Kelvin-Voigt solid rho u_tt = (G + mu_s d/dt)(u_yy + u_y / y), both checked
  by finite differences in y and t.  A planted sign error in the interface
  stress (G + i omega mu_s) fails the stress check.
* The extrapolation reproduces, at a tight tolerance, the linear fields a
  planar interface admits exactly (eta = r across the interface r = L_s,
  eta = z along it), zeroes eta outside its box, and fails with the
  normal's sign flipped.
* The driver's two Poisson routes agree over 60 steps; its first step has
  dt = 0 (the sample timer starts at its limit); dt then follows the three
  bounds of :216-220.
"""
import numpy as np
import pytest

from oracle.drivers import PeriodicSoftSlabCase, soft_slab_theory
from oracle.pde_extrapolation import StaticPDEExtrapolation

L_F = L_S = 0.1
OMEGA = 2 * np.pi
V_WALL = 1.0
SHEAR = 2 * V_WALL / OMEGA / (2 * (L_F + L_S))
NU = SHEAR * OMEGA * L_F**2 / 10
G = NU * SHEAR * OMEGA / 0.25


def _theory(**kw):
    return soft_slab_theory(L_F, L_S, SHEAR, OMEGA, G, V_WALL, 1.0, NU, **kw)


def test_theory_wall_and_interface():
    temporal, Y, (fluid, solid, disp) = _theory()
    for t in (0.0, 0.13, 0.4, 0.77):
        assert abs(fluid(L_S + L_F, t) - V_WALL * np.cos(OMEGA * t)) < 1e-12
        assert abs(temporal(t)[-1] - V_WALL * np.cos(OMEGA * t)) < 1e-12
        assert abs(fluid(L_S, t) - solid(L_S, t)) < 1e-12 * max(1.0, abs(fluid(L_S, t)))
    # shear stress matched at the interface
    h, mu = 1e-6, NU
    for t in (0.05, 0.3, 0.61):
        tf = mu * (fluid(L_S + h, t) - fluid(L_S - h, t)) / (2 * h)
        ts = (G * (disp(L_S + h, t) - disp(L_S - h, t)) + mu * (solid(L_S + h, t)
                                                               - solid(L_S - h, t))) / (2 * h)
        assert abs(tf - ts) < 1e-6 * max(abs(tf), 1e-3)
    assert np.all(np.isfinite(temporal(0.3))) and Y[0] > 0 and Y[-1] == pytest.approx(L_S + L_F)


def _residual_fluid(fluid):
    y = np.linspace(L_S + 0.01, L_S + L_F - 0.01, 81)
    h, k, t = 1e-4, 1e-5, 0.37
    f = lambda yy, tt: fluid(yy, tt)  # noqa: E731
    vt = (f(y, t + k) - f(y, t - k)) / (2 * k)
    vyy = (f(y + h, t) - 2 * f(y, t) + f(y - h, t)) / h**2
    vy = (f(y + h, t) - f(y - h, t)) / (2 * h)
    return np.max(np.abs(vt - NU * (vyy + vy / y))) / np.max(np.abs(vt))


def _residual_solid(disp, mu):
    y = np.linspace(0.01, L_S - 0.01, 81)
    h, k, t = 1e-4, 1e-4, 0.21
    u = lambda yy, tt: disp(yy, tt)  # noqa: E731
    utt = (u(y, t + k) - 2 * u(y, t) + u(y, t - k)) / k**2
    lap = lambda tt: ((u(y + h, tt) - 2 * u(y, tt) + u(y - h, tt)) / h**2  # noqa: E731
                      + (u(y + h, tt) - u(y - h, tt)) / (2 * h) / y)
    rhs = G * lap(t) + mu * (lap(t + k) - lap(t - k)) / (2 * k)
    return np.max(np.abs(utt - rhs)) / np.max(np.abs(utt))


def test_theory_satisfies_both_equations():
    _, _, (fluid, solid, disp) = _theory()
    assert _residual_fluid(fluid) < 1e-3
    assert _residual_solid(disp, NU) < 1e-3


def test_planted_interface_stress_sign_fails():
    from scipy.special import jv, yv

    # the same system with (G + i omega mu_s) in the interface stress
    lam1 = np.sqrt(1j * OMEGA / NU)
    lam2 = OMEGA / np.sqrt(-1j * OMEGA * NU + G)
    L = L_S + L_F
    k8 = lam2 * jv(1, lam2 * L_S) / (NU * lam1) * (G + NU * 1j * OMEGA)
    M = np.array([[jv(0, lam1 * L), yv(0, lam1 * L), 0],
                  [jv(0, lam1 * L_S), yv(0, lam1 * L_S), 1j * OMEGA * jv(0, lam2 * L_S)],
                  [jv(1, lam1 * L_S), yv(1, lam1 * L_S), -k8]], dtype=complex)
    A, B, C = np.linalg.solve(M, np.array([0.5, 0, 0], dtype=complex))
    fluid = lambda y, t: 2 * np.real((A * jv(0, lam1 * y) + B * yv(0, lam1 * y))  # noqa: E731
                                     * np.exp(-1j * OMEGA * t))
    disp = lambda y, t: 2 * np.real(C * jv(0, lam2 * y) * np.exp(-1j * OMEGA * t))  # noqa: E731
    solid = lambda y, t: 2 * np.real(-1j * OMEGA * C * jv(0, lam2 * y)  # noqa: E731
                                     * np.exp(-1j * OMEGA * t))
    h, t = 1e-6, 0.3
    tf = NU * (fluid(L_S + h, t) - fluid(L_S - h, t)) / (2 * h)
    ts = (G * (disp(L_S + h, t) - disp(L_S - h, t)) + NU * (solid(L_S + h, t)
                                                           - solid(L_S - h, t))) / (2 * h)
    assert abs(tf - ts) > 1e-3 * max(abs(tf), 1e-3)


def _slab_grid(nr=96, nz=12, dx=1.0 / 192):
    r = (np.arange(nr) + 0.5) * dx
    z = (np.arange(nz) + 0.5) * dx
    Z, R = np.meshgrid(z, r)
    return Z, R, dx


@pytest.mark.parametrize("flip", [False, True])
def test_extrapolation_planar_interface_exact(flip):
    Z, R, dx = _slab_grid()
    nr, nz = R.shape
    L_s = 0.2
    band = np.sqrt(2) * dx + 3 * dx
    phi = L_s - R
    inside = phi > 0
    ex = StaticPDEExtrapolation(dx, nr, nz, extrap_tol=1e-13, extrap_band=band)
    errs = []
    for exact in (0.3 + 2.0 * R, 0.1 + 0.0 * R + Z):
        eta = exact * inside
        ex.extrapolate(eta, -phi if flip else phi)
        zone = (phi <= 0) & (phi > -band + 1e-12)
        zone[:, :2] = zone[:, -2:] = False  # the box keeps its z margins
        errs.append(np.max(np.abs(eta[zone] - exact[zone])))
        if not flip:
            # cells beyond the box are zeroed (eta *= 0, :228)
            assert np.all(eta[ex.r_end - 2 + 1:, :] == 0)
    if flip:
        assert max(errs) > 1e-2
    else:
        assert max(errs) < 1e-10, errs


def test_driver_routes_dt_and_ghosts():
    a = PeriodicSoftSlabCase(solver="exact")
    b = PeriodicSoftSlabCase(solver="eig")
    for _ in range(60):
        a.step()
        b.step()
    for k in ("vorticity", "psi", "u_z", "eta1", "eta2"):
        x, y = getattr(a, k), getattr(b, k)
        assert np.linalg.norm(x - y) <= 1e-9 * np.linalg.norm(x), k
    assert a.dts[0] == 0.0
    dx = a.dx
    bound = min(a.CFL * dx / np.sqrt(a.G / a.rho_f), 0.9 * dx**2 / 4 / a.nu)
    assert a.dts[1] == pytest.approx(bound, rel=1e-15)
    g = a.g
    assert np.array_equal(a.psi[:, :g], a.psi[:, -2 * g:-g])
    assert a.nr == 256 and a.nz == 8
    assert len(a.samples) == 1 and a.samples[0][0] == 0.0
