"""The allocation-free force sum (axs_penalization_force_sum_noalloc_f64)
against the workspace entry point axs_penalization_force_sum_f64, through
the C ABI: bit for bit on two calls in a row, the second on a smaller
input (the library-owned partials keep the first call's contents), and an
oversize n is refused before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sum(lib, fn, R, chi, uz, U, n, partials=None):
    import torch

    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    if partials is None:
        st = getattr(lib, fn)(ptr(R), ptr(chi), ptr(uz), ctypes.c_double(U), ctypes.c_int64(n),
                              ptr(out), s)
    else:
        st = getattr(lib, fn)(ptr(R), ptr(chi), ptr(uz), ctypes.c_double(U), ctypes.c_int64(n),
                              ptr(partials), ptr(out), s)
    assert st == 0, lib.axs_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()[0]


def test_noalloc_force_sum_matches_workspace_entry_point():
    import torch

    from pyaxisymflow_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(7)
    for nr, nz in ((1024, 2048), (37, 53)):  # second call smaller
        n = nr * nz
        R = torch.from_numpy(np.repeat(rng.uniform(0, 1, (nr, 1)), nz, axis=1)).cuda()
        chi = torch.from_numpy(rng.uniform(0, 1, (nr, nz))).cuda()
        uz = torch.from_numpy(rng.standard_normal((nr, nz))).cuda()
        partials = torch.empty(lib.axs_reduce_partials_size(n), dtype=torch.float64,
                               device="cuda")
        ref = _sum(lib, "axs_penalization_force_sum_f64", R, chi, uz, 0.3, n, partials)
        got = _sum(lib, "axs_penalization_force_sum_noalloc_f64", R, chi, uz, 0.3, n)
        assert got == ref, (nr, nz, got, ref)
        terms = R.cpu().numpy() * chi.cpu().numpy() * (uz.cpu().numpy() - 0.3)
        host = float(np.sum(terms))
        assert abs(got - host) <= 1e-13 * float(np.sum(np.abs(terms)))
    # a size beyond the limit is refused with a message
    st = lib.axs_penalization_force_sum_noalloc_f64(
        ctypes.c_void_p(R.data_ptr()), ctypes.c_void_p(chi.data_ptr()),
        ctypes.c_void_p(uz.data_ptr()), ctypes.c_double(0.0), ctypes.c_int64((1 << 40) + 1),
        ctypes.c_void_p(R.data_ptr()), None)
    assert st == _lib.AXS_ERR_INVALID and b"2^40" in lib.axs_last_error()
