"""Pins the C restatement of the least-squares extrapolation (oracle/lsq.c):
the reference's own Gauss known-answer test, the compiled reference header
(oracle/_ref, built from the reference tree where it exists), a committed
golden fixture made from that build, exactness on linear fields and the
start-code table.  CPU only."""
import os

import numpy as np
import pytest

from oracle import lsq
from oracle.kernels import smooth_Heaviside

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "lsq_soft_sphere_nz64.npz")


def test_gauss_reference_kat():
    """core/src/lstsq/test_least_squares.cpp:30-61 (tolerance 1e-14)."""
    A = np.array([[2.56, 0.86, 4.2, 1.964, 1.284],
                  [0.86, 0.32, 1.4, 0.666, 0.45],
                  [4.2, 1.4, 7.0, 3.22, 2.1]])
    sol = lsq.gauss_elimination_3x2(A)
    assert np.abs(sol - np.array([[0.7, 0.2, 0.0], [0.3, 0.6, 0.0]])).max() <= 1e-14


def test_start_code_table():
    """ExtrapolationConfig.hpp:131-158 for the 3x3 window: code =
    int(2 p' / n) with p' = p rounded up to even when p > 4; start = code - 3.
    Spot values: all neighbours known -> window centred (-1); only left
    column known -> window to the left (-3); three right cells -> (+1)."""
    assert lsq.start_code(8, 8) == -1
    assert lsq.start_code(0, 3) == -3
    assert lsq.start_code(6, 3) == 1
    assert lsq.start_code(5, 3) == 1   # 5 -> 6 (even) -> int(12/3) - 3
    assert lsq.start_code(3, 2) == 0   # int(6/2) - 3
    for p in range(0, 17):
        for n in range(1, 9):
            pe = (2 * int(np.ceil(0.5 * p))) if p > 4 else p
            assert lsq.start_code(p, n) == int(2.0 * pe / n) - 3


def _soft_sphere(nz, seed=1, noise=0.01):
    dx = 1.0 / nz
    nr = nz // 2
    z = np.linspace(dx / 2, 1 - dx / 2, nz)
    r = np.linspace(dx / 2, 0.5 - dx / 2, nr)
    Z, R = np.meshgrid(z, r)
    moll = 4 * dx
    ez = moll + 4 * dx
    rng = np.random.default_rng(seed)
    phi = -np.sqrt((Z - 0.5) ** 2 + R**2) + 0.15
    chi = 0 * Z
    smooth_Heaviside(chi, phi, moll)
    inside = chi > 0.5
    eta1 = Z + noise * rng.standard_normal(Z.shape)
    eta2 = R.copy()
    phid = np.zeros((nz, nz))
    e1 = np.zeros((nz, nz))
    e2 = np.zeros((nz, nz))
    phid[nr:] = -phi
    e1[nr:] = inside * eta1
    e2[nr:] = inside * eta2
    phid[:nr] = -np.flip(phi, 0)
    e1[:nr] = np.flip(e1[nr:], 0)
    e2[:nr] = -np.flip(e2[nr:], 0)
    cf = np.asarray(phid < 0, dtype=np.int16)
    tf = np.asarray(phid < ez, dtype=np.int16)
    return cf, tf, e1, e2, z


def test_linear_field_reproduced_exactly():
    """First-order LS reproduces a linear eta exactly wherever it extrapolates."""
    cf, tf, _, _, z = _soft_sphere(64)
    X, Y = np.meshgrid(z, z)
    ex = np.where(cf == 1, 0.3 + 1.1 * X - 0.7 * Y, 0.0)
    ey = np.where(cf == 1, -0.2 + 0.4 * X + 2.0 * Y, 0.0)
    c0 = cf.copy()
    n = lsq.extrapolate_using_least_squares_till_first_order(cf, tf, ex, ey, z, z)
    assert n > 3
    band = (cf == 1) & (c0 == 0)
    assert band.sum() > 100
    np.testing.assert_allclose(ex[band], (0.3 + 1.1 * X - 0.7 * Y)[band], atol=1e-12)
    np.testing.assert_allclose(ey[band], (-0.2 + 0.4 * X + 2.0 * Y)[band], atol=1e-12)
    assert np.array_equal(cf, tf)  # every target cell reached


def test_matches_golden_fixture():
    """Fixture written by tests/golden/make_golden.py from the compiled
    reference header: flags and eta agree bit for bit."""
    g = np.load(GOLDEN)
    cf, tf, ex, ey = (g[k].copy() for k in ("cur_in", "tgt", "ex_in", "ey_in"))
    lsq.extrapolate_using_least_squares_till_first_order(cf, tf, ex, ey, g["gx"], g["gy"])
    assert np.array_equal(cf, g["cur_out"])
    assert np.array_equal(ex, g["ex_out"]) and np.array_equal(ey, g["ey_out"])


@pytest.mark.skipif(not lsq.ref_available(), reason="reference build (oracle/_ref) absent")
@pytest.mark.parametrize("nz,seed", [(64, 1), (128, 2), (256, 3)])
def test_matches_compiled_reference(nz, seed):
    cf, tf, e1, e2, z = _soft_sphere(nz, seed)
    a = [cf.copy(), tf.copy(), e1.copy(), e2.copy()]
    b = [cf.copy(), tf.copy(), e1.copy(), e2.copy()]
    lsq.extrapolate_using_least_squares_till_first_order(*a, z, z)
    lsq.ref_extrapolate_using_least_squares_till_first_order(*b, z, z)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    A = np.array([[2.0, 1.0, 0.5, 1.0, 2.0], [1.0, 3.0, 0.2, 0.0, 1.0], [0.5, 0.2, 4.0, 1.0, 1.0]])
    assert np.array_equal(lsq.gauss_elimination_3x2(A), lsq.gauss_elimination_3x2(A, lsq.ref_lib()))
