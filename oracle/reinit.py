"""Signed-distance reinitialisation of a level set in a narrow band (row A15).
TEST ORACLE ONLY.

The reference calls scikit-fmm 2022.8.15, ``skfmm.distance(phi, dx=dx,
narrow=band)`` (examples/SoftSphereInTaylorGreenVortex/
soft_sphere_in_taylor_green_vortex.py:159-162), which is neither vendored
nor installed here.  This restates its published algorithm:

1. Interface cells (phi == 0, or a sign change towards an axis neighbour)
   are frozen at the distance interpolated linearly to the crossing along
   each axis, d_a = dx phi_i / (phi_i - phi_nb) (nearest crossing per axis),
   combined as |d| = 1 / sqrt(sum_a 1/d_a^2), with the sign of phi.
2. Every other cell takes the upwind eikonal update |grad d| = 1 from its
   smaller axis neighbours: second order (coefficient 9/4, value
   (4 d1 - d2)/3) when the next cell in that direction is known and not
   larger, first order otherwise; the quadratic's larger root is taken, and
   where it is not above both axes' upwind values the update uses the axis
   with the smaller one alone.
3. Instead of marching in order of distance, the update is iterated over
   all cells (Jacobi sweeps, each replacing the iterate) until it stops
   changing.  At the fixed point every cell satisfies its upwind update
   from its converged neighbours, the equations fast marching solves in
   order of distance.
4. Cells with |d| > narrow are masked (returned as the input phi).

Accuracy (measured; tests/test_oracle_reinit.py asserts it):
* the marching is second order: with the frozen ring set to the exact
  distance of a sphere, the max error in the band away from the ring falls
  by 2^1.7-2^1.9 per halving of dx (a first-order update falls by 2^0.9);
* the frozen ring itself is first order at oblique interfaces: a cell with
  a crossing along one axis only takes that axis's intercept, which
  overestimates the distance by up to a factor 1/cos(angle).  That is the
  published initialisation of scikit-fmm's distance marcher, so the max
  error of the whole result (ring included) halves per halving of dx.

Parity with scikit-fmm itself is unpinned (the package is absent): fast
marching and Jacobi sweeps reach the same discrete equations but not the
same bits.
"""
import numpy as np

BIG = 1e300


def interface_distance(phi, dx):
    """Step 1: frozen mask and their signed distances (nan elsewhere)."""
    nr, nz = phi.shape
    inv = np.zeros_like(phi)        # sum of 1/d_a^2
    border = np.zeros(phi.shape, dtype=bool)
    for axis in (0, 1):
        dmin = np.full(phi.shape, np.inf)
        for shift in (-1, 1):
            nb = np.full(phi.shape, np.nan)
            if axis == 0:
                if shift == -1:
                    nb[1:, :] = phi[:-1, :]
                else:
                    nb[:-1, :] = phi[1:, :]
            else:
                if shift == -1:
                    nb[:, 1:] = phi[:, :-1]
                else:
                    nb[:, :-1] = phi[:, 1:]
            with np.errstate(invalid="ignore", divide="ignore"):
                cross = np.isfinite(nb) & (phi * nb < 0)
                d = np.where(cross, dx * phi / (phi - nb), np.inf)
            dmin = np.minimum(dmin, d)
        has = np.isfinite(dmin)
        border |= has
        with np.errstate(divide="ignore"):
            inv = inv + np.where(has, 1.0 / dmin / dmin, 0.0)
    zero = phi == 0.0
    dist = np.full(phi.shape, np.nan)
    with np.errstate(divide="ignore"):
        mag = np.sqrt(1.0 / inv)
    dist[border] = np.where(phi[border] < 0, -mag[border], mag[border])
    dist[zero] = 0.0
    return border | zero, dist


def _shift(a, axis, k, fill):
    out = np.full(a.shape, fill)
    if axis == 0:
        if k > 0:
            out[:-k, :] = a[k:, :]
        else:
            out[-k:, :] = a[:k, :]
    else:
        if k > 0:
            out[:, :-k] = a[:, k:]
        else:
            out[:, -k:] = a[:, :k]
    return out


def eikonal_update(D, side, dx):
    """Step 2 for every cell: D holds signed distances of known cells and
    BIG for unknown ones; neighbour values are taken relative to the cell's
    own side (side = +1 / -1).  Per axis, v1 is the smaller known neighbour
    and v2 the next cell beyond it when known and not larger; the axis term
    is second order, (9/4)(x - (4 v1 - v2)/3)^2 / dx^2, with v2 and first
    order, (x - v1)^2 / dx^2, without.  x is the larger root of
    sum(terms) = 1.  When both axes enter and that root is not above both
    axes' upwind values (the upwind condition fails), x comes from the axis
    with the smaller upwind value alone: its value + 2 dx / 3 (second order)
    or + dx (first order).  Returns the new distance magnitude per cell (BIG
    where no neighbour is known)."""
    known = D < BIG
    a = np.zeros_like(D)
    b = np.zeros_like(D)
    c = np.zeros_like(D)
    tps, steps, used = [], [], []
    idx2 = 1.0 / (dx * dx)
    for axis in (0, 1):
        v1 = np.full_like(D, BIG)
        v2 = np.full_like(D, BIG)
        for s in (-1, 1):
            k1 = _shift(known, axis, s, False)
            k2 = _shift(known, axis, 2 * s, False)
            n1 = np.where(k1, side * _shift(D, axis, s, 0.0), BIG)
            n2 = np.where(k2, side * _shift(D, axis, 2 * s, 0.0), BIG)
            better = n1 < v1
            v2 = np.where(better, np.where(n2 <= n1, n2, BIG), v2)
            v1 = np.where(better, n1, v1)
        second = (v1 < BIG) & (v2 < BIG)
        first = (v1 < BIG) & ~second
        w1 = np.where(v1 < BIG, v1, 0.0)
        w2 = np.where(second, v2, 0.0)
        tp = np.where(second, (4.0 * w1 - w2) / 3.0, w1)
        aa = 9.0 / 4.0
        a = a + np.where(second, idx2 * aa, np.where(first, idx2, 0.0))
        b = b - np.where(second, idx2 * 2 * aa * tp, np.where(first, idx2 * 2 * w1, 0.0))
        c = c + np.where(second, idx2 * aa * tp * tp, np.where(first, idx2 * w1 * w1, 0.0))
        tps.append(np.where(second | first, tp, BIG))
        steps.append(np.where(second, (2.0 / 3.0) * dx, dx))
        used.append(second | first)
    c = c - 1.0
    det = b * b - 4.0 * a * c
    with np.errstate(invalid="ignore", divide="ignore"):
        x = (-b + np.sqrt(np.maximum(det, 0.0))) / (2.0 * a)
    both = used[0] & used[1]
    upwind_fail = both & ((det < 0) | (x < tps[0]) | (x < tps[1]))
    one = np.where(tps[0] <= tps[1], tps[0] + steps[0], tps[1] + steps[1])
    x = np.where(upwind_fail, one, x)
    return np.where(a > 0, x, BIG)


def march(frozen, dist, phi, dx, iterations):
    """Step 3: Jacobi sweeps of the eikonal update over every cell that is
    not frozen; each sweep replaces the iterate (values may move either
    way: the second-order update is not monotone in its neighbours, and
    keeping only decreases would freeze stale values).  `dist` holds the
    frozen cells' signed distances."""
    side = np.where(phi < 0, -1.0, 1.0)
    D = np.where(frozen, np.nan_to_num(dist), BIG)
    for _ in range(iterations):
        m = eikonal_update(D, side, dx)
        D = np.where(frozen | (m >= BIG), D, side * m)
    return D


def distance(phi, dx, narrow, iterations=None):
    """Signed distance of phi's zero level set within |d| <= narrow.
    Returns (d, masked): d equals phi where masked."""
    frozen, dist = interface_distance(phi, dx)
    if iterations is None:
        iterations = 2 * int(np.ceil(narrow / dx)) + 8
    D = march(frozen, dist, phi, dx, iterations)
    d = np.where(frozen, dist, D)
    masked = ~(np.abs(d) <= narrow)
    return np.where(masked, phi, d), masked
