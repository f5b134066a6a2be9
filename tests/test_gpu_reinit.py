"""GPU level-set reinitialisation (row A15) against its CPU restatement
(oracle/reinit.py, itself pinned by closed forms in test_oracle_reinit.py).
Same FP64 operations in the same order, so the bar is bit-exact; also the
closed forms directly on the GPU result (aligned plane exact; a sphere
within 0.25 dx, converging at first order from the frozen ring's axis
intercepts; mask = |d| > narrow)."""
import numpy as np
import pytest

from oracle.reinit import distance as cpu_distance

pytestmark = pytest.mark.gpu


def _grid(nz):
    dx = 1.0 / nz
    nr = nz // 2
    z = np.linspace(dx / 2, 1 - dx / 2, nz)
    r = np.linspace(dx / 2, 0.5 - dx / 2, nr)
    Z, R = np.meshgrid(z, r)
    return dx, Z, R


def _fields(nz):
    dx, Z, R = _grid(nz)
    sphere = 0.15 - np.sqrt((Z - 0.5) ** 2 + R**2)
    return dx, Z, R, {
        "sphere_tanh": 0.3 * np.tanh(5.0 * sphere),
        "sphere_exact": sphere,
        "plane": 3.0 * (Z - 0.4321),
        "two_blobs": np.maximum(0.1 - np.sqrt((Z - 0.3) ** 2 + R**2),
                                0.08 - np.sqrt((Z - 0.7) ** 2 + (R - 0.2) ** 2)),
    }


@pytest.mark.parametrize("nz", [64, 256])
@pytest.mark.parametrize("name", ["sphere_tanh", "sphere_exact", "plane", "two_blobs"])
def test_reinit_matches_oracle_bitwise(nz, name):
    from pyaxisymflow_amd.levelset import distance

    dx, Z, R, fields = _fields(nz)
    phi = fields[name]
    band = 8 * dx
    print(f"reinit {name} nz={nz}", flush=True)
    got = distance(phi, dx=dx, narrow=band)
    want, masked = cpu_distance(phi, dx, band)
    assert np.array_equal(got.mask, masked)
    assert np.array_equal(np.asarray(got), want)


def test_reinit_closed_forms():
    from pyaxisymflow_amd.levelset import distance

    dx, Z, R, fields = _fields(128)
    band = 6 * dx
    got = distance(fields["plane"], dx=dx, narrow=band)
    exact = Z - 0.4321
    inb = ~got.mask
    np.testing.assert_allclose(np.asarray(got)[inb], exact[inb], atol=1e-14)
    assert np.array_equal(got.mask, np.abs(exact) > band + 1e-12)
    sphere = fields["sphere_exact"]
    got = distance(fields["sphere_tanh"], dx=dx, narrow=8 * dx)
    inb = ~got.mask
    assert np.abs(np.asarray(got) - sphere)[inb].max() < 0.25 * dx
    # masked entries keep phi (what the driver restores)
    assert np.array_equal(np.asarray(got)[got.mask], fields["sphere_tanh"][got.mask])
    errs = []
    for nz in (128, 256, 512):
        dx, Z, R, fields = _fields(nz)
        got = distance(fields["sphere_tanh"], dx=dx, narrow=8 * dx)
        errs.append(np.abs(np.asarray(got) - fields["sphere_exact"])[~got.mask].max())
    orders = [np.log2(errs[k] / errs[k + 1]) for k in range(2)]
    assert all(0.8 < o < 1.3 for o in orders), (errs, orders)


def test_reinit_rejects_unsupported_options():
    from pyaxisymflow_amd.levelset import distance

    phi = np.zeros((8, 16))
    with pytest.raises(NotImplementedError):
        distance(phi, dx=0.1, order=1)
    with pytest.raises(NotImplementedError):
        distance(phi, dx=0.1, periodic=True)
