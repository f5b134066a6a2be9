"""Pins the level-set reinitialisation oracle (oracle/reinit.py, the
scikit-fmm `distance` restatement; scikit-fmm itself is absent, so parity
with it is unpinned).  Closed forms: grid-aligned planes are reproduced
exactly (the upwind updates are exact on linear fields); the frozen ring of
an oblique plane is exact where both axes cross and the axis intercept
where one does; the marching converges at second order from an exact ring
(a planted first-order update fails that); the whole result converges at
first order (the ring's intercepts), |grad d| = 1 in the band, odd
symmetry, and the narrow-band mask.  CPU only."""
import numpy as np
import pytest

import oracle.reinit as RI
from oracle.reinit import distance, interface_distance, march


def _grid(nz):
    dx = 1.0 / nz
    nr = nz // 2
    z = np.linspace(dx / 2, 1 - dx / 2, nz)
    r = np.linspace(dx / 2, 0.5 - dx / 2, nr)
    Z, R = np.meshgrid(z, r)
    return dx, Z, R


@pytest.mark.parametrize("z0", [0.4, 0.4321])
def test_aligned_plane_exact(z0):
    dx, Z, R = _grid(64)
    exact = Z - z0
    phi = 3.0 * exact  # same zero set, slope 3: not a distance
    d, masked = distance(phi, dx, 6 * dx)
    band = ~masked
    assert band.sum() > 10 * Z.shape[0]
    np.testing.assert_allclose(d[band], exact[band], atol=1e-14)
    # a nonlinear monotone phi moves only the interpolated crossing
    d2, _ = distance(3.0 * np.tanh(2.0 * exact), dx, 6 * dx)
    np.testing.assert_allclose(d2[band], exact[band], atol=1e-4 * dx)
    assert np.array_equal(masked, np.abs(exact) > 6 * dx + 1e-12)
    assert np.array_equal(d[masked], phi[masked])


def test_interface_cells_plane_interpolation():
    dx, Z, R = _grid(32)
    phi = 2.0 * (R - 0.2031)
    frozen, dist = interface_distance(phi, dx)
    rows = np.where(frozen.any(axis=1))[0]
    assert len(rows) == 2  # the two rows straddling r = 0.2031
    np.testing.assert_allclose(dist[frozen], (R - 0.2031)[frozen], atol=1e-14)


def _sphere(nz):
    dx, Z, R = _grid(nz)
    exact = 0.15 - np.sqrt((Z - 0.5) ** 2 + R**2)
    return dx, exact, 0.3 * np.tanh(5.0 * exact)


def _orders(errs):
    return [float(np.log2(errs[k] / errs[k + 1])) for k in range(len(errs) - 1)]


def _marching_errors():
    """Max error of the marching alone: the frozen ring set to the exact
    sphere distance, error over the band |d| <= 0.05 off the ring."""
    errs = []
    for nz in (128, 256, 512):
        dx, exact, phi = _sphere(nz)
        frozen, _ = interface_distance(phi, dx)
        D = march(frozen, np.where(frozen, exact, np.nan), phi, dx,
                  2 * int(np.ceil(0.0625 / dx)) + 8)
        sel = ~frozen & (np.abs(exact) <= 0.05)
        errs.append(np.abs(D - exact)[sel].max())
    return errs


def test_marching_is_second_order():
    errs = _marching_errors()
    orders = _orders(errs)
    assert min(orders) > 1.65, (errs, orders)
    assert errs[-1] < 1e-5


def test_planted_first_order_update_fails_the_order_check(monkeypatch):
    """The order check discriminates: an eikonal update that ignores the
    second neighbour (first order) converges at order < 1.2."""
    monkeypatch.setattr(RI, "eikonal_update", _first_order_update)
    orders = _orders(_marching_errors())
    assert max(orders) < 1.2, orders


def _first_order_update(D, side, dx):
    """The first-order upwind update (the planted mistake)."""
    BIG = RI.BIG
    known = D < BIG
    a = np.zeros_like(D)
    b = np.zeros_like(D)
    c = np.full_like(D, -1.0)
    for axis in (0, 1):
        v1 = np.full_like(D, BIG)
        for s in (-1, 1):
            n1 = np.where(RI._shift(known, axis, s, False), side * RI._shift(D, axis, s, 0.0), BIG)
            v1 = np.minimum(v1, n1)
        has = v1 < BIG
        w = np.where(has, v1, 0.0)
        a += np.where(has, 1.0, 0.0) / dx**2
        b -= np.where(has, 2 * w, 0.0) / dx**2
        c += np.where(has, w * w, 0.0) / dx**2
    with np.errstate(invalid="ignore", divide="ignore"):
        x = (-b + np.sqrt(np.maximum(b * b - 4 * a * c, 0.0))) / (2 * a)
    return np.where(a > 0, x, BIG)


@pytest.mark.parametrize("angle", [0.4, 0.7, np.pi / 2 - 0.4])
def test_oblique_plane_frozen_ring(angle):
    """A plane at an angle: a ring cell with crossings along both axes gets
    the exact distance 1/sqrt(1/dz^2 + 1/dr^2); one with a crossing along
    one axis only gets that axis's intercept (scikit-fmm's initialisation),
    which is larger than the distance."""
    dx, Z, R = _grid(64)
    nz_, nr_ = np.cos(angle), np.sin(angle)
    exact = nz_ * (Z - 0.5) + nr_ * (R - 0.2) + 0.1234 * dx
    frozen, dist = interface_distance(2.0 * exact, dx)
    cz = np.zeros_like(Z, dtype=bool)
    cr = np.zeros_like(Z, dtype=bool)
    sgn = np.sign(exact)
    cz[:, 1:] |= sgn[:, 1:] != sgn[:, :-1]
    cz[:, :-1] |= sgn[:, 1:] != sgn[:, :-1]
    cr[1:, :] |= sgn[1:, :] != sgn[:-1, :]
    cr[:-1, :] |= sgn[1:, :] != sgn[:-1, :]
    both = frozen & cz & cr
    only_z = frozen & cz & ~cr
    only_r = frozen & cr & ~cz
    assert both.sum() > 10 and (only_z | only_r).sum() > 10
    np.testing.assert_allclose(dist[both], exact[both], atol=1e-14)
    np.testing.assert_allclose(dist[only_z], exact[only_z] / nz_, atol=1e-14)
    np.testing.assert_allclose(dist[only_r], exact[only_r] / nr_, atol=1e-14)


def test_sphere_result_first_order_ring():
    """The whole result (ring included): max error about 0.2 dx, halving per
    halving of dx (order 0.8-1.3), from the ring's intercepts."""
    errs = []
    for nz in (128, 256, 512):
        dx, exact, phi = _sphere(nz)
        d, masked = distance(phi, dx, 8 * dx)
        errs.append(np.abs(d - exact)[~masked].max())
        assert errs[-1] < 0.25 * dx
    orders = _orders(errs)
    assert all(0.8 < o < 1.3 for o in orders), (errs, orders)


@pytest.mark.parametrize("nz", [64, 128])
def test_sphere_eikonal_and_symmetry(nz):
    dx, Z, R = _grid(nz)
    exact = 0.15 - np.sqrt((Z - 0.5) ** 2 + R**2)
    phi = 0.3 * np.tanh(5.0 * exact)
    band_w = 8 * dx
    d, masked = distance(phi, dx, band_w)
    inb = ~masked
    gz = np.gradient(d, dx, axis=1)
    gr = np.gradient(d, dx, axis=0)
    inner = inb & (np.abs(exact) < band_w - 2 * dx) & (np.abs(exact) > 2 * dx)
    assert np.median(np.abs(np.hypot(gz, gr) - 1)[inner]) < 0.05
    # odd symmetry
    d2, m2 = distance(-phi, dx, band_w)
    assert np.array_equal(m2, masked)
    np.testing.assert_allclose(d2[inb], -d[inb], atol=1e-14)


def test_converged():
    dx, Z, R = _grid(64)
    phi = 0.3 * np.tanh(5.0 * (0.15 - np.sqrt((Z - 0.5) ** 2 + R**2)))
    a, _ = distance(phi, dx, 8 * dx)
    b, _ = distance(phi, dx, 8 * dx, iterations=200)
    assert np.array_equal(a, b)
