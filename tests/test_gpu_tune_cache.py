"""The column pass's layout record (AXS_TUNE_CACHE): the first process that
times the layouts for a grid writes its choice, and a later process reads it
instead of timing again, so both solve with the same layout and give the same
bits.  A record the library did not write (a valid layout other than the
searched one) is taken as it stands.  Grid 512 x 512 (above the size where
the search runs), drop-in Stokes solver, seeded right-hand side."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SOLVE = r"""
import sys
import numpy as np
sys.path.insert(0, {root!r})
from pyaxisymflow_amd.kernels.FastDiagonalisationStokesSolver import FastDiagonalisationStokesSolver
nr = nz = 512
rng = np.random.default_rng(3)
rhs = rng.standard_normal((nr, nz))
psi = np.zeros_like(rhs)
FastDiagonalisationStokesSolver(nr, nz, 1.0 / nz).solve(psi, rhs)
np.save({out!r}, psi)
"""


def _solve(tmp_path, cache, tag):
    out = str(tmp_path / f"psi_{tag}.npy")
    env = dict(os.environ, AXS_TUNE_CACHE=str(cache))
    env.pop("AXS_TRIDIAG_LK", None)
    env.pop("AXS_NO_TUNE", None)
    r = subprocess.run([sys.executable, "-c", SOLVE.format(root=ROOT, out=out)], env=env,
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    return np.load(out)


def test_layout_record_written_then_reused(tmp_path):
    cache = tmp_path / "tridiag.txt"
    print("tune cache: first process times and records", flush=True)
    a = _solve(tmp_path, cache, "a")
    lines = cache.read_text().split("\n")
    rec = [ln.split() for ln in lines if ln.strip()]
    assert len(rec) == 1 and rec[0][1:4] == ["512", "512", "0"], rec
    L, K, G = map(int, rec[0][4:])
    assert L in (4, 8, 16) and K >= 1 and G in (0, 1)
    print(f"tune cache: recorded L={L} K={K} G={G}; second process reads it", flush=True)
    b = _solve(tmp_path, cache, "b")
    assert np.array_equal(a, b)
    assert cache.read_text().count("\n") == 1  # read, not searched and appended again


def test_foreign_layout_record_is_used(tmp_path):
    cache = tmp_path / "tridiag.txt"
    a = _solve(tmp_path, cache, "a")
    arch = cache.read_text().split()[0]
    rec = cache.read_text().split()
    L0 = int(rec[4])
    # another usable layout: segments of 8 rows (4 when the search took 8),
    # two modes per workgroup, coefficients in LDS
    L1 = 4 if L0 == 8 else 8
    cache.write_text(f"{arch} 512 512 0 {L1} 2 0\n")
    print(f"tune cache: forced L={L1} K=2 (searched L={L0})", flush=True)
    b = _solve(tmp_path, cache, "b")
    # a different segmentation rounds differently but solves the same system
    assert np.linalg.norm(b - a) <= 1e-12 * np.linalg.norm(a)
    assert cache.read_text() == f"{arch} 512 512 0 {L1} 2 0\n"
