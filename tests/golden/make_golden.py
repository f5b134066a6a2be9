"""Writes tests/golden/*.npz fixtures from the reference's own compiled C++
(oracle/_ref/libref_lsq.so, built by `make -C oracle` from the reference
tree).  Run in the build container:  python tests/golden/make_golden.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import lsq  # noqa: E402
from test_oracle_lsq import _soft_sphere  # noqa: E402


def main():
    cf, tf, e1, e2, z = _soft_sphere(64, seed=7)
    c, ex, ey = cf.copy(), e1.copy(), e2.copy()
    lsq.ref_extrapolate_using_least_squares_till_first_order(c, tf, ex, ey, z, z)
    np.savez_compressed(
        os.path.join(HERE, "lsq_soft_sphere_nz64.npz"),
        cur_in=cf, tgt=tf, ex_in=e1, ey_in=e2, gx=z, gy=z, cur_out=c, ex_out=ex, ey_out=ey,
    )


if __name__ == "__main__":
    main()
