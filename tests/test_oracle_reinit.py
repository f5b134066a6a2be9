"""Pins the level-set reinitialisation oracle (oracle/reinit.py, the
scikit-fmm `distance` restatement; scikit-fmm itself is absent, so parity
with it is unpinned).  Closed forms: grid-aligned planes are reproduced
exactly (the upwind updates are exact on linear fields), a sphere to O(dx),
|grad d| = 1 in the band, odd symmetry, and the narrow-band mask.  CPU only."""
import numpy as np
import pytest

from oracle.reinit import distance, interface_distance


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


@pytest.mark.parametrize("nz", [64, 128])
def test_sphere_distance_and_eikonal(nz):
    dx, Z, R = _grid(nz)
    exact = 0.15 - np.sqrt((Z - 0.5) ** 2 + R**2)
    phi = 0.3 * np.tanh(5.0 * exact)
    band_w = 8 * dx
    d, masked = distance(phi, dx, band_w)
    inb = ~masked
    assert np.abs(d - exact)[inb].max() < 0.5 * dx
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
