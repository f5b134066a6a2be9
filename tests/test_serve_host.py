"""Host side of the serving form: the synthetic request set (CPU only)."""
import numpy as np


def test_request_set_sizes_are_supported():
    """nr = clip(round(lognormal(ln 256, 0.5)), 128, 1024), not rounded to
    transform-friendly sizes: the set holds sizes with large prime factors."""
    from pyaxisymflow_amd.serve import request_set

    rs = request_set(200, seed=3, rate=10.0)
    for r in rs:
        assert 128 <= r["nr"] <= 1024 and r["grid_size_z"] == 2 * r["nr"]
        assert 0.1 <= r["G"] <= 10.0

    def largest_prime_factor(n):
        p, f = 2, 1
        while n > 1:
            while n % p == 0:
                n //= p
                f = p
            p += 1
        return f

    assert max(largest_prime_factor(r["nr"]) for r in rs) > 5
    arr = np.array([r["arrival"] for r in rs])
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
