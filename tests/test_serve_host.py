"""Host side of the serving form: the synthetic request set (CPU only)."""
import numpy as np


def test_request_set_sizes_are_supported():
    from pyaxisymflow_amd.serve import request_set, smooth_size

    rs = request_set(200, seed=3, rate=10.0)
    for r in rs:
        assert 128 <= r["nr"] <= 1024 and r["grid_size_z"] == 2 * r["nr"]
        assert smooth_size(r["nr"], 128, 1024) == r["nr"]
        assert 0.1 <= r["G"] <= 10.0
    arr = np.array([r["arrival"] for r in rs])
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
