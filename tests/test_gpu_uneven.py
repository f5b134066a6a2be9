"""The batched form over uneven grids: cases of several grids advanced by one
launch per pass (axs_rigid_batch_create with differing nr, nz, dx).

* With the radial solve of both pinned to one layout (segments of 8 rows,
  as many modes per workgroup as fit 1024 threads), every case's
  vorticity and psi equal those of the same cases stepped as one-grid
  batches, bit for bit; the force record differs only by the summation
  order of the per-tile partials (the completion reduction runs with a
  different workgroup size).
* The velocity_max tile height picked at set-up (8, 16 or 32 rows) leaves
  every field and the drag unchanged, bit for bit.
* Each case also matches the oracle's restatement of the driver loop
  (flow_past_sphere.py:107-188) within the FP64 bar.
* FP32 uneven batches keep the drag within 1e-4 of FP64.
* Unsupported batches (a moving body, nz > 1024) are refused.
"""
import numpy as np
import pytest

from oracle.drivers import RigidBodyCase

pytestmark = pytest.mark.gpu

# (nz, nr), z in [0, 1], r in [0, 0.5]; nz = 1024 runs the one-wave row passes
GRIDS = [(128, 64), (192, 96), (96, 48), (1024, 512)]


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def uneven_cases(seeds_per_grid):
    from pyaxisymflow_amd import workloads

    cases = []
    for (nz, nr), seeds in zip(GRIDS, seeds_per_grid):
        for s in seeds:
            sp = workloads.case_spec(s)
            sp.update(nz=nz, nr=nr)
            _, chi = workloads.case_inputs(sp, nz, nr)
            a = workloads.rigid_case_args(sp, chi)
            a.update(nz=nz, nr=nr, z_len=1.0, r_len=0.5)
            cases.append(a)
    return cases


def grid_args(case):
    return {k: v for k, v in case.items() if k not in ("nz", "nr", "z_len", "r_len")}


def test_uneven_equals_one_grid_batches(monkeypatch):
    from pyaxisymflow_amd.stepper import RigidBatchStepper, UnevenBatchStepper

    # interleaved on purpose: the stepper groups the cases by grid
    cases = uneven_cases([[0, 3], [1, 4, 6], [2], [5, 7]])
    order = [0, 2, 6, 1, 5, 3, 7, 4]
    mixed = [cases[q] for q in order]
    monkeypatch.setenv("AXS_TRIDIAG_LK", "8,1")  # segments of 8 rows (K is per group)
    uneven = UnevenBatchStepper(mixed)
    uneven.step(30)
    for g, (nz, nr) in enumerate(GRIDS):
        P = -(-nr // 8)
        monkeypatch.setenv("AXS_TRIDIAG_LK", f"8,{max(1, 1024 // P)}")
        idx = [q for q, c in enumerate(mixed) if (c["nz"], c["nr"]) == (nz, nr)]
        one = RigidBatchStepper(nz, nr, 1.0, [grid_args(mixed[q]) for q in idx], r_len=0.5)
        one.step(30)
        for k, q in enumerate(idx):
            print(f"grid {nz}x{nr} case {q}", flush=True)
            assert uneven.vorticity(q).shape == (nr, nz)
            assert np.array_equal(uneven.vorticity(q), one.vorticity(k))
            assert np.array_equal(uneven.psi(q), one.psi(k))
            assert uneven.time(q) == one.time(k) and uneven.steps_done(q) == 30
            np.testing.assert_allclose(uneven.drag_history(q), one.drag_history(k), rtol=1e-13,
                                       atol=1e-15)


@pytest.mark.parametrize("tr", [8, 32])
def test_uneven_vmax_tile_height_is_exact(monkeypatch, tr):
    """The velocity_max tile height the set-up search picks (8, 16 or 32
    rows) changes only the tiling of an exact maximum: vorticity, psi and
    drag equal those of 16-row tiles bit for bit."""
    from pyaxisymflow_amd.stepper import UnevenBatchStepper

    cases = uneven_cases([[0], [1], [2]])
    monkeypatch.setenv("AXS_TRIDIAG_LK", "8")
    runs = []
    for rows in (16, tr):
        monkeypatch.setenv("AXS_VMAX_TR", str(rows))
        b = UnevenBatchStepper(cases)
        b.step(20)
        runs.append(b)
    for q in range(len(cases)):
        assert np.array_equal(runs[0].vorticity(q), runs[1].vorticity(q))
        assert np.array_equal(runs[0].psi(q), runs[1].psi(q))
        assert np.array_equal(runs[0].drag_history(q), runs[1].drag_history(q))


def test_uneven_cases_match_oracle():
    from pyaxisymflow_amd.stepper import UnevenBatchStepper

    cases = uneven_cases([[0], [1], [2]])
    gpu = UnevenBatchStepper(cases)
    gpu.step(25)
    for q, c in enumerate(cases):
        cpu = RigidBodyCase(c["nz"], c["nr"], c["z_len"], c["nu"], c["brink_lam"], c["char_func"],
                            U_0=c["U_0"], T_ramp=c["T_ramp"], r_sph=c["r_sph"], solver="exact")
        for _ in range(25):
            cpu.step()
        assert rel_l2(gpu.vorticity(q), cpu.vorticity) < 1e-10
        assert rel_l2(gpu.psi(q), cpu.psi) < 1e-10
        np.testing.assert_allclose(gpu.drag_history(q), np.array(cpu.drag), rtol=1e-10,
                                   atol=1e-12)


def test_uneven_fp32_drag_error():
    from pyaxisymflow_amd.stepper import UnevenBatchStepper

    cases = uneven_cases([[0, 5], [1, 7], [2, 3]])
    f64 = UnevenBatchStepper(cases, precision="f64")
    f32 = UnevenBatchStepper(cases, precision="f32")
    f64.step(150)
    f32.step(150)
    for q in range(len(cases)):
        d64, d32 = f64.drag_history(q), f32.drag_history(q)
        err = np.max(np.abs(d32 - d64)) / np.max(np.abs(d64))
        print(f"case {q}: max rel drag err {err:.2e}", flush=True)
        assert err < 1e-4


def test_uneven_refusals():
    from pyaxisymflow_amd.stepper import UnevenBatchStepper

    cases = uneven_cases([[0], [1], []])
    moving = dict(z0=0.5, r0=0.0, aspect=1.0, radius=0.1, disp=0.01, moll_zone=0.02)
    bad = [dict(c, moving_body=moving, oscillation=(0.1, 6.0)) for c in cases]
    with pytest.raises((ValueError, RuntimeError)):
        UnevenBatchStepper(bad)
    big = dict(cases[0], nz=2048, nr=1024, char_func=np.zeros((1024, 2048)), vorticity=None)
    with pytest.raises((ValueError, RuntimeError)):
        UnevenBatchStepper([cases[1], big])


def test_uneven_fp32_against_the_fp64_oracle():
    """An FP32 uneven batch against the FP64 oracle's restatement of the
    driver loop (exact Poisson route), 100 steps: relative L2 of the
    penalised velocity and psi, and the drag relative to its maximum, all
    below 1e-4."""
    from pyaxisymflow_amd.stepper import UnevenBatchStepper

    cases = uneven_cases([[0], [1], [2]])
    gpu = UnevenBatchStepper(cases, precision="f32")
    gpu.step(100)
    for q, c in enumerate(cases):
        cpu = RigidBodyCase(c["nz"], c["nr"], c["z_len"], c["nu"], c["brink_lam"], c["char_func"],
                            U_0=c["U_0"], T_ramp=c["T_ramp"], r_sph=c["r_sph"], solver="exact")
        for _ in range(100):
            cpu.step()
        uz, ur = gpu.velocity(q)
        err_u = rel_l2(np.hypot(uz, ur), np.hypot(cpu.u_z, cpu.u_r))
        err_psi = rel_l2(gpu.psi(q), cpu.psi)
        d = gpu.drag_history(q)
        ref = np.array(cpu.drag)
        err_d = np.abs(d - ref).max() / np.abs(ref).max()
        print(f"f32 case {q} ({c['nz']}x{c['nr']}): u {err_u:.2e} psi {err_psi:.2e} "
              f"drag {err_d:.2e}", flush=True)
        assert err_u < 1e-4 and err_psi < 1e-4 and err_d < 1e-4
