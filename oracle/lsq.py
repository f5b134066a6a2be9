"""ctypes access to the oracle's C restatement of the least-squares eta
extrapolation (lsq.c) and to the reference's own header compiled into
_ref/libref_lsq.so (ref_lsq_driver.cpp).  TEST ORACLE ONLY.

Both are built by ``make -C oracle``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LSQ = os.path.join(_HERE, "_lsq.so")
_REF = os.path.join(_HERE, "_ref", "libref_lsq.so")

_p16 = np.ctypeslib.ndpointer(np.int16, flags="C_CONTIGUOUS")
_pf = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def _load(path, prefix):
    lib = ctypes.CDLL(path)
    f = getattr(lib, prefix + "extrapolate_first_order")
    f.argtypes = [_p16, _p16, _pf, _pf, _pf, _pf, ctypes.c_int, ctypes.c_int]
    g = getattr(lib, prefix + "gauss_elimination_3x2")
    g.argtypes = [_pf, _pf]
    s = getattr(lib, prefix + "start_code")
    s.argtypes = [ctypes.c_short, ctypes.c_short]
    s.restype = ctypes.c_short
    return lib


_lib = None
_ref = None


def oracle_lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LSQ):
            raise RuntimeError("oracle/_lsq.so missing: run `make -C oracle`")
        _lib = _load(_LSQ, "oracle_")
    return _lib


def ref_available():
    return os.path.exists(_REF)


def ref_lib():
    global _ref
    if _ref is None:
        _ref = _load(_REF, "ref_")
    return _ref


def _check(current_flag, target_flag, eta_x, eta_y, grid_x, grid_y):
    for a, dt in ((current_flag, np.int16), (target_flag, np.int16),
                  (eta_x, np.float64), (eta_y, np.float64),
                  (grid_x, np.float64), (grid_y, np.float64)):
        if a.dtype != dt or not a.flags.c_contiguous:
            raise TypeError("wrong dtype or layout")
    n0, n1 = current_flag.shape
    if grid_x.shape[0] < n1 or grid_y.shape[0] < n0:
        raise ValueError("grid vectors too short")
    return n0, n1


def extrapolate_using_least_squares_till_first_order(
    current_flag, target_flag, eta_x, eta_y, grid_x, grid_y
):
    """Oracle restatement of core/src/extrapolate_using_least_squares.hpp:450-467
    (in place on current_flag, eta_x, eta_y).  Returns the sweep count."""
    n0, n1 = _check(current_flag, target_flag, eta_x, eta_y, grid_x, grid_y)
    return oracle_lib().oracle_extrapolate_first_order(
        current_flag, target_flag, eta_x, eta_y, grid_x, grid_y, n0, n1
    )


def ref_extrapolate_using_least_squares_till_first_order(
    current_flag, target_flag, eta_x, eta_y, grid_x, grid_y
):
    """The reference header itself (compiled from /root/reference)."""
    n0, n1 = _check(current_flag, target_flag, eta_x, eta_y, grid_x, grid_y)
    ref_lib().ref_extrapolate_first_order(
        current_flag, target_flag, eta_x, eta_y, grid_x, grid_y, n0, n1
    )


def gauss_elimination_3x2(A, lib=None):
    """Returns the (2, 3) solution of the 3x(3+2) augmented system A."""
    lib = lib or oracle_lib()
    A = np.ascontiguousarray(A, dtype=np.float64).copy()
    sol = np.zeros(18)
    fn = lib.oracle_gauss_elimination_3x2 if lib is oracle_lib() else lib.ref_gauss_elimination_3x2
    fn(A.reshape(-1), sol)
    return sol.reshape(2, 9)[:, :3]


def start_code(positional_sum, total_sum, lib=None):
    lib = lib or oracle_lib()
    fn = lib.oracle_start_code if lib is oracle_lib() else lib.ref_start_code
    return int(fn(positional_sum, total_sum))
