"""Pins the numpy oracle (oracle/kernels.py) with checks that do not restate
its formulas: the reference's own test, closed forms, exactness on
polynomials, discrete null vectors and conservation.  CPU only."""
import numpy as np
import pytest

from oracle import kernels as K
from oracle.drivers import grid


def _grid(nz=64, nr=32, z_len=1.0):
    return grid(nz, nr, z_len)


# ---------------------------------------------------------------- A5
@pytest.mark.parametrize("n", [16])
def test_brinkmann_reference_kat(n):
    """tests/test_kernels/test_brinkmann_penalize.py:6-26 (the reference's
    only unit test): lam=2, dt=3, chi=4, U=(1,2), zero grid velocity."""
    lam, dt = 2.0, 3.0
    chi = 4 * np.ones((n, n))
    U_z, U_r = 1.0, 2.0
    gz = np.zeros((n, n))
    gr = np.zeros((n, n))
    pz = np.zeros((n, n))
    pr = np.zeros((n, n))
    K.brinkmann_penalize(lam, dt, chi, U_z, U_r, gz, gr, pz, pr)
    np.testing.assert_allclose(pz, lam * dt * U_z * chi / (1 + lam * dt * chi))
    np.testing.assert_allclose(pr, lam * dt * U_r * chi / (1 + lam * dt * chi))
    np.testing.assert_allclose(pz, 24 / 25)  # 6*4*1/(1+24)


def test_brinkmann_limits():
    """chi=0 leaves u; lam*dt*chi -> inf drives u to U."""
    u = np.linspace(-1, 1, 16).reshape(4, 4)
    out = np.zeros_like(u)
    out2 = np.zeros_like(u)
    K.brinkmann_penalize(1e4, 1e-3, np.zeros_like(u), 5.0, 6.0, u, u, out, out2)
    assert np.array_equal(out, u)
    K.brinkmann_penalize(1e30, 1.0, np.ones_like(u), 5.0, 6.0, u, u, out, out2)
    np.testing.assert_allclose(out, 5.0)
    np.testing.assert_allclose(out2, 6.0)


# ---------------------------------------------------------------- A6
def test_heaviside_closed_form():
    w = 0.1
    phi = np.array([[-2 * w, -w, -w / 2, 0.0, w / 2, w, 2 * w]])
    H = np.zeros_like(phi)
    K.smooth_Heaviside(H, phi, w)
    expect = [0.0, 0.0, 0.5 * (0.5 - 1 / np.pi), 0.5, 0.5 * (1.5 + 1 / np.pi), 1.0, 1.0]
    np.testing.assert_allclose(H[0], expect, atol=1e-15)


# ---------------------------------------------------------------- A1
def test_kill_taper_constant_field_unit_domain():
    """For z_max = 1 and width 3 the weights are sin(0), sin(pi/4), sin(pi/2)
    at the left end (mirrored at the right), times the value of column 2
    (resp. nz-3)."""
    nz, nr = 64, 32
    dx, z, r, Z, R = _grid(nz, nr, 1.0)
    w = np.full((nr, nz), 2.0)
    K.kill_boundary_vorticity_sine_z(w, Z, 3, dx)
    h = np.sqrt(2.0)
    np.testing.assert_allclose(w[:, :3], np.tile([0.0, h, 2.0], (nr, 1)), atol=1e-14)
    np.testing.assert_allclose(w[:, -3:], np.tile([2.0, h, 0.0], (nr, 1)), atol=1e-14)
    assert np.all(w[:, 3:-3] == 2.0)


def test_kill_taper_literal_one_quirk():
    """The reference hard-codes the domain end as 1: on z in [0, 2] the
    right-hand weights are sin(pi (1 - z - dx/2) / (4 dx)), not a taper to
    zero at z = 2.  The two differ by the phase pi (1 - 2) / (4 dx); with
    nz = 260 (dx = 1/130) that is -32.5 pi, so the literal weights are
    -cos of the taper's angle: -1 at the last column where a taper to z = 2
    would give 0.  The oracle keeps the literal form."""
    nz, nr = 260, 130
    dx, z, r, Z, R = _grid(nz, nr, 2.0)
    w = np.ones((nr, nz))
    K.kill_boundary_vorticity_sine_z(w, Z, 3, dx)
    theta = np.pi * (2.0 - z[-3:] - 0.5 * dx) / (4 * dx)  # taper angle to z = 2
    np.testing.assert_allclose(w[5, -3:], -np.cos(theta), atol=1e-12)
    np.testing.assert_allclose(w[5, -1], -1.0, atol=1e-12)
    # the first row below the taper is untouched, the axis row is not part of the z taper
    assert w[5, -4] == 1.0


def test_kill_r_zeroes_axis_row_and_tapers_top():
    nz, nr = 32, 16
    dx, z, r, Z, R = grid(nz, nr, 1.0, r_len=0.5)
    w = np.arange(nr * nz, dtype=float).reshape(nr, nz)
    src = w[-3].copy()
    K.kill_boundary_vorticity_sine_r(w, R, 3, dx)
    assert np.all(w[0] == 0)
    for q, i in enumerate(range(nr - 3, nr)):
        np.testing.assert_allclose(
            w[i], np.sin(np.pi * (1 - r[i] - 0.5 * dx) / (4 * dx)) * src, rtol=1e-14
        )


# ---------------------------------------------------------------- A3 / A4
def test_velocity_exact_on_quadratics():
    """Central and one-sided 2nd-order differences are exact on quadratics:
    psi = a r^2 + b r z + c z^2 gives u_z r = 2 a r + b z, -u_r r = b r + 2 c z."""
    dx, z, r, Z, R = _grid(48, 24)
    a, b, c = 0.7, -1.3, 0.4
    psi = a * R**2 + b * R * Z + c * Z**2
    uz = np.zeros_like(psi)
    ur = np.zeros_like(psi)
    K.compute_velocity_from_psi_unb(uz, ur, psi, R, dx)
    np.testing.assert_allclose(uz * R, 2 * a * R + b * Z, atol=1e-11)
    np.testing.assert_allclose(-ur * R, b * R + 2 * c * Z, atol=1e-11)


def test_vorticity_exact_on_quadratics_and_border_untouched():
    dx, z, r, Z, R = _grid(48, 24)
    uz = 0.3 * R**2 - 0.2 * Z * R
    ur = 0.5 * Z**2 + 0.1 * Z * R
    w = np.full_like(uz, 7.0)
    K.compute_vorticity_from_velocity_unb(w, uz, ur, dx)
    expect = (1.0 * Z + 0.1 * R) - (0.6 * R - 0.2 * Z)  # d(u_r)/dz - d(u_z)/dr
    np.testing.assert_allclose(w[1:-1, 1:-1], expect[1:-1, 1:-1], atol=1e-11)
    assert np.all(w[0] == 7) and np.all(w[-1] == 7) and np.all(w[:, 0] == 7) and np.all(w[:, -1] == 7)


# ---------------------------------------------------------------- A8
def test_diffusion_discrete_null_vector():
    """omega = r (a + b z) is a discrete null vector of
    L = lap + (1/r) d/dr - 1/r^2 on the interior, so RK2 leaves it fixed."""
    dx, z, r, Z, R = _grid(64, 32)
    w = R * (0.3 + 1.7 * Z)
    w0 = w.copy()
    tmp = np.zeros_like(w)
    K.diffusion_RK2_unb(w, tmp, R, 1e-2, 1e-3, dx)
    np.testing.assert_allclose(w, w0, rtol=0, atol=1e-12)


def test_diffusion_r_cubed_closed_form():
    """omega = r^3: the discrete operator gives exactly 8 r + dx^2 / r, so the
    first stage is tmp = r^3 + (nu dt / 2)(8 r + dx^2/r) on the interior."""
    dx, z, r, Z, R = _grid(64, 32)
    w = R**3
    tmp = np.zeros_like(w)
    nu, dt = 0.02, 1e-3
    w_in = w.copy()
    K.diffusion_RK2_unb(w, tmp, R, nu, dt, dx)
    Lw = 8 * R + dx**2 / R
    np.testing.assert_allclose(tmp[1:-1, 1:-1], (w_in + 0.5 * nu * dt * Lw)[1:-1, 1:-1], rtol=1e-12)
    assert np.array_equal(tmp[0], w_in[0])


# ---------------------------------------------------------------- A7 / A13
def _linear_setup(uz_val, ur_val=0.0, nz=64, nr=32):
    dx, z, r, Z, R = _grid(nz, nr)
    uz = np.full((nr, nz), uz_val)
    ur = np.full((nr, nz), ur_val)
    return dx, Z, R, uz, ur


@pytest.mark.parametrize("u", [0.8, -0.6])
def test_eno3_nonconservative_linear_exact(u):
    """eta1 = a + b z with uniform u_z: one Euler step gives a + b z - dt u b
    on the advected cells (both upwind branches reproduce the face value)."""
    dx, Z, R, uz, ur = _linear_setup(u)
    e1 = 0.2 + 1.5 * Z
    e2 = R.copy()
    e10, e20 = e1.copy(), e2.copy()
    dt = 1e-3
    K.advect_refmap_via_eno3(e1, e2, uz, ur, dt, dx)
    nr, nz = e1.shape
    np.testing.assert_allclose(e1[:nr - 2, 2:-2], (e10 - dt * u * 1.5)[:nr - 2, 2:-2], atol=1e-13)
    # never-advected cells keep their values (ghost layers of width 2)
    assert np.array_equal(e1[nr - 2:], e10[nr - 2:])
    assert np.array_equal(e1[:, :2], e10[:, :2]) and np.array_equal(e1[:, -2:], e10[:, -2:])
    # eta2 = r with zero u_r is unchanged (u_z d/dz r = 0)
    np.testing.assert_allclose(e2, e20, atol=1e-14)


def test_eno3_nonconservative_constant_unchanged():
    dx, Z, R, uz, ur = _linear_setup(0.5)
    rng = np.random.default_rng(3)
    uz = rng.standard_normal(uz.shape)
    ur = rng.standard_normal(ur.shape)
    e1 = np.full_like(uz, 3.25)
    e2 = np.zeros_like(uz)
    K.advect_refmap_via_eno3(e1, e2, uz, ur, 1e-3, dx)
    np.testing.assert_allclose(e1, 3.25, atol=1e-13)


def test_eno3_conservative_telescopes_in_z():
    """With u_r = 0 the conservative update of a field supported well inside
    the domain sums to zero along every row (face fluxes cancel)."""
    dx, z, r, Z, R = _grid(64, 32)
    rng = np.random.default_rng(5)
    uz = rng.uniform(-1, 1, Z.shape)
    ur = np.zeros_like(uz)
    w = np.zeros_like(uz)
    w[4:20, 10:50] = rng.standard_normal((16, 40))
    w0 = w.copy()
    K.advect_vorticity_via_eno3(w, uz, ur, 1e-3, dx)
    np.testing.assert_allclose(w.sum(axis=1), w0.sum(axis=1), atol=1e-12)
    assert not np.allclose(w, w0)


def test_eno3_conservative_uniform_flow_linear_field():
    """f = a + b z, uniform u_z: flux difference = u b dx, so f decreases by
    dt u b on the advected cells."""
    dx, Z, R, uz, ur = _linear_setup(0.7)
    w = 1.0 + 2.0 * Z
    w0 = w.copy()
    dt = 2e-3
    K.advect_vorticity_via_eno3(w, uz, ur, dt, dx)
    nr = w.shape[0]
    np.testing.assert_allclose(w[:nr - 2, 2:-2], (w0 - dt * 0.7 * 2.0)[:nr - 2, 2:-2], atol=1e-12)


def test_eno3_axis_mirror_parities():
    """A unit delta on the first row with u_r = -1 everywhere.  Across the
    axis the mirror gives u_r(-1) = +1 (odd) and omega(-1) = -1 (odd); both
    faces take their second ENO3 branch, so the closed form is
    omega_new = 1 - (5/6) dt/dx.  An even omega, or an even u_r, would give
    1 - (1/6) dt/dx instead."""
    dx, z, r, Z, R = _grid(32, 16)
    uz = np.zeros_like(Z)
    ur = np.full_like(Z, -1.0)
    w = np.zeros_like(Z)
    w[0, 16] = 1.0
    dt = 1e-3
    K.advect_vorticity_via_eno3(w, uz, ur, dt, dx)
    np.testing.assert_allclose(w[0, 16], 1.0 - (5.0 / 6.0) * dt / dx, rtol=1e-14)


# ---------------------------------------------------------------- A11 / A12
def test_solid_sigma_identity_map_zero_where_written():
    dx, z, r, Z, R = _grid(32, 16)
    s11, s12, s22 = (np.zeros_like(Z) for _ in range(3))
    g = [np.zeros_like(Z) for _ in range(4)]
    K.solid_sigma(s11, s12, s22, 2.0, dx, Z.copy(), R.copy(), *g)
    # interior: eta_z = (1, 0), eta_r = (0, 1) -> sigma = 0
    np.testing.assert_allclose(s11[1:-1, 1:-1], 0, atol=1e-12)
    np.testing.assert_allclose(s12[1:-1, 1:-1], 0, atol=1e-12)
    # the never-written eta_r entries of the top row stay 0 -> sigma11 = -G/2 * eta1z^2
    np.testing.assert_allclose(s11[-1, 1:-1], -0.5 * 2.0 * 1.0, rtol=1e-12)
    np.testing.assert_allclose(s22, -s11)


def test_solid_sigma_linear_shear_and_hoop_terms():
    """eta1 = z + g r, eta2 = r: sigma12 = -G g, sigma11 = G g^2 / 2 on the
    interior; then tau_z = sigma12 / r, tau_r = sigma22 / r (hoop terms), and
    omega gains dt * d(tau_r)/dz - dt * d(tau_z)/dr = dt * G g / r^2 ... in
    discrete form: dt * (-(tau_z[i+1] - tau_z[i-1]) / (2 dx))."""
    dx, z, r, Z, R = _grid(32, 16)
    G, gam = 1.5, 0.3
    s11, s12, s22 = (np.zeros_like(Z) for _ in range(3))
    g = [np.zeros_like(Z) for _ in range(4)]
    K.solid_sigma(s11, s12, s22, G, dx, Z + gam * R, R.copy(), *g)
    np.testing.assert_allclose(s12[1:-1, 1:-1], -G * gam, rtol=1e-12)
    np.testing.assert_allclose(s11[1:-1, 1:-1], 0.5 * G * gam**2, rtol=1e-10)
    # constant interior stress field: use it directly for div_tau
    t11 = np.full_like(Z, 0.5 * G * gam**2)
    t12 = np.full_like(Z, -G * gam)
    t22 = -t11
    w = np.zeros_like(Z)
    tz, tr = np.zeros_like(Z), np.zeros_like(Z)
    dt = 1e-2
    K.update_vorticity_from_solid_stress(w, tz, tr, t11, t12, t22, R, dt, dx)
    np.testing.assert_allclose(tz[1:-1, 1:-1], (t12 / R)[1:-1, 1:-1], rtol=1e-12)
    np.testing.assert_allclose(tr[1:-1, 1:-1], (t22 / R)[1:-1, 1:-1], rtol=1e-12)
    # rows 2..nr-3 see only interior tau (rows 1 and nr-2 read the row-0 /
    # untouched border values)
    expect = -dt * (t12[3:-1, 2:-2] / R[3:-1, 2:-2] - t12[1:-3, 2:-2] / R[1:-3, 2:-2]) / (2 * dx)
    np.testing.assert_allclose(w[2:-2, 2:-2], expect, rtol=1e-10)


# ---------------------------------------------------------------- A9 / A10
def test_force_and_max_reductions():
    dx, z, r, Z, R = _grid(32, 16)
    chi = (Z < 0.5).astype(float)
    uz = np.full_like(Z, 2.0)
    F_pen, F_un = K.compute_force_on_body(R, chi, 1.0, 3.0, uz, 0.5, 2.0, 0.1, 0.3)
    np.testing.assert_allclose(F_pen, 3.0 * 1.5 * (R * chi).sum(), rtol=1e-14)
    np.testing.assert_allclose(F_un, 0.3 * 2.0 / 0.1)
    ur = -np.abs(Z)
    assert K.max_abs_velocity_sum(uz, ur) == 2.0 + np.abs(Z).max()
