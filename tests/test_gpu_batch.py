"""The batched rigid-body form: many cases of one grid stepped by the same
launches (axs_rigid_batch_create).

* Every case of a batch equals the same case stepped alone, bit for bit
  (same kernels, same per-case reductions), so the single-case parity
  against the oracle (test_gpu_stepper.py) carries over to each case.
* One case of a batch also matches the oracle's restatement of the driver
  loop directly (flow_past_sphere.py:107-188).
* FP32 batched drag stays within 1e-4 of FP64 on the case set's bodies
  (sphere, prolate spheroid, torus) after 200 steps.
"""
import numpy as np
import pytest

from oracle.drivers import RigidBodyCase

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make_cases(seeds, nz, nr):
    from pyaxisymflow_amd import workloads

    specs, args = [], []
    for s in seeds:
        sp = workloads.case_spec(s)
        sp.update(nz=nz, nr=nr)
        _, chi = workloads.case_inputs(sp, nz, nr)
        specs.append(sp)
        args.append(workloads.rigid_case_args(sp, chi))
    return specs, args


def test_batch_cases_equal_single_runs():
    from pyaxisymflow_amd.stepper import RigidBatchStepper, RigidBodyStepper

    nz, nr = 128, 64
    specs, args = make_cases([0, 1, 2, 5], nz, nr)
    bodies = {s["body"] for s in specs}
    print("bodies", sorted(bodies), "nu", [round(s["nu"], 6) for s in specs], flush=True)
    batch = RigidBatchStepper(nz, nr, 1.0, args, r_len=0.5)
    batch.step(25)
    for q, a in enumerate(args):
        one = RigidBodyStepper(nz, nr, 1.0, r_len=0.5, **a)
        one.step(25)
        assert np.array_equal(batch.vorticity(q), one.vorticity())
        assert np.array_equal(batch.drag_history(q), one.drag_history())
        assert batch.time(q) == one.time()
    # cases advance their own dt (different viscosities and bodies)
    assert len({batch.dt(q) for q in range(len(args))}) > 1


def test_batch_case_matches_oracle():
    from pyaxisymflow_amd.stepper import RigidBatchStepper

    nz, nr = 128, 64
    specs, args = make_cases([3, 4], nz, nr)
    batch = RigidBatchStepper(nz, nr, 1.0, args, r_len=0.5)
    batch.step(40)
    for q, a in enumerate(args):
        cpu = RigidBodyCase(nz, nr, 1.0, a["nu"], a["brink_lam"], a["char_func"], U_0=a["U_0"],
                            T_ramp=a["T_ramp"], r_sph=a["r_sph"], solver="exact")
        for _ in range(40):
            cpu.step()
        assert rel_l2(batch.vorticity(q), cpu.vorticity) < 1e-10
        np.testing.assert_allclose(batch.drag_history(q), np.array(cpu.drag), rtol=1e-10,
                                   atol=1e-12)


def test_batched_fp32_drag_error():
    from pyaxisymflow_amd.stepper import RigidBatchStepper

    nz, nr = 256, 128
    specs, args = make_cases([0, 1, 2, 3, 5, 7], nz, nr)
    f64 = RigidBatchStepper(nz, nr, 1.0, args, r_len=0.5, precision="f64")
    f32 = RigidBatchStepper(nz, nr, 1.0, args, r_len=0.5, precision="f32")
    f64.step(200)
    f32.step(200)
    for q, sp in enumerate(specs):
        d64, d32 = f64.drag_history(q), f32.drag_history(q)
        err = np.max(np.abs(d32 - d64)) / np.max(np.abs(d64))
        print(f"case {sp['seed']} {sp['body']}: Cd {d64[-1]:.6f} max rel err {err:.2e}", flush=True)
        assert err < 1e-4
