"""The C ABI: the library loads without a GPU, exports every function that
include/axisym.h declares, and the Python binding declares the same set.
No compute calls (CPU only)."""
import ctypes
import os
import re

import pytest

from pyaxisymflow_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "axisym.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(axs_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert len(names) >= 20
    assert "axs_stokes_solve_f64" in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_version_and_error_without_gpu():
    lib = _lib.load()
    assert lib.axs_version().startswith(b"axisym-hip")
    # argument validation happens before any HIP call
    st = lib.axs_brinkmann_penalize_f64(1.0, 1.0, None, 0.0, 0.0, None, None, None, None, 4, None)
    assert st == _lib.AXS_ERR_INVALID
    assert b"null" in lib.axs_last_error()
    plan = ctypes.c_void_p()
    st = lib.axs_stokes_plan_create(2, 16, 0.1, 0, ctypes.byref(plan))  # nr < 3
    assert st == _lib.AXS_ERR_INVALID and not plan.value
    with pytest.raises(ValueError):
        _lib.check(st, "plan")


def test_dropin_rejects_wrong_dtype_before_gpu():
    import numpy as np

    from pyaxisymflow_amd.kernels.brinkmann_penalize import brinkmann_penalize

    a = np.zeros((4, 4), dtype=np.float32)
    with pytest.raises(TypeError):
        brinkmann_penalize(1.0, 1.0, a, 0.0, 0.0, a, a, a, a)
