"""Drop-in for the level-set reinitialisation the soft-body driver calls:
``skfmm.distance(phi, dx=dx, narrow=band)`` (scikit-fmm 2022.8.15) at
examples/SoftSphereInTaylorGreenVortex/soft_sphere_in_taylor_green_vortex.py:159-162.

``distance`` keeps scikit-fmm's signature for the arguments that call site
uses and returns what it returns: a ``numpy.ma.MaskedArray`` whose mask is
True outside the band (|d| > narrow), where the driver restores the old
phi.  The distance itself comes from the HIP kernels in axs_reinit.hip
(linear interface distances, then Jacobi sweeps of the second-order upwind
eikonal update).
"""
import ctypes

import numpy as np

from pyaxisymflow_amd import _lib


def distance(phi, dx=1.0, self_test=False, order=2, narrow=0.0, periodic=False,
             iterations=None):
    """Signed distance to the zero level set of ``phi`` (2D, C order),
    computed within ``narrow`` of the interface (narrow <= 0: everywhere).

    ``dx`` must be a scalar (the reference grids are square).  Only the
    reference's options are supported: order 2, no periodic axes."""
    if order != 2:
        raise NotImplementedError("only order=2 (the reference's) is supported")
    if periodic not in (False, None) and np.any(periodic):
        raise NotImplementedError("periodic boundaries are not supported")
    if not np.isscalar(dx):
        raise NotImplementedError("dx must be one scalar (square grid)")
    a = np.ascontiguousarray(np.asarray(phi, dtype=np.float64))
    if a.ndim != 2:
        raise ValueError("phi must be 2D")
    nr, nz = a.shape
    if narrow is None or narrow <= 0:
        band = float(np.inf)
        iters = 2 * (nr + nz) + 8 if iterations is None else int(iterations)
    else:
        band = float(narrow)
        iters = -1 if iterations is None else int(iterations)
    st = _lib.Staged()
    out = np.empty_like(a)
    mask = np.empty(a.shape, dtype=np.uint8)
    _lib.call(
        "axs_reinit_distance_f64",
        st.ptr(a, "phi"),
        st.ptr(out, "out", out=True, copy_in=False),
        st.ptr(mask, "masked", dtype=np.uint8, out=True, copy_in=False),
        nr, nz, float(dx), band if np.isfinite(band) else 1e300, iters,
        _lib.stream_handle(),
    )
    st.finish()
    return np.ma.MaskedArray(out, mask=mask.astype(bool))
