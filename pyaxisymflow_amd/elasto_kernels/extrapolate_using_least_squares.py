"""Drop-in for pyaxisymflow/elasto_kernels/extrapolate_using_least_squares.py:8-38."""
import numpy as np

from pyaxisymflow_amd import _lib
from pyaxisymflow_amd.core.extrapolate_using_least_squares import (
    extrapolate_using_least_squares_till_first_order as extrapolate_using_least_squares,
)


def extrapolate_eta_using_least_squares(
    inp_phi,
    phi_thresh_lower_bound: float,
    phi_thresh_upper_bound: float,
    inp_eta_X,
    inp_eta_Y,
    inp_x,
    inp_y,
):
    """Extrapolates inp_eta_X, inp_eta_Y from {phi < lower} to
    {phi < upper} (extrapolate_using_least_squares.py:8-38)."""
    if _lib.is_device_tensor(inp_phi):
        t = _lib.torch()
        current_flag = (inp_phi < phi_thresh_lower_bound).to(t.int16)
        target_flag = (inp_phi < phi_thresh_upper_bound).to(t.int16)
    else:
        current_flag = np.asarray(inp_phi < phi_thresh_lower_bound, dtype=np.int16)
        target_flag = np.asarray(inp_phi < phi_thresh_upper_bound, dtype=np.int16)
    extrapolate_using_least_squares(
        current_flag, target_flag, inp_eta_X, inp_eta_Y, inp_x, inp_y
    )
