"""Checkpoint and resume (state / load_state / save / load): a stepper built
with the same parameters and loaded with another's state continues bit for
bit as that one does.

* Rigid sphere, 256 x 128, with a ramped free stream and the drag record:
  31 steps (an odd count: the vorticity pair ends on its second buffer
  before the run's completion brings it home), a checkpoint through an .npz
  file, then 20 more steps in both: vorticity, psi, scalars and the drag
  history equal.
* Soft sphere (single stepper, 128 x 64): 17 steps, checkpoint, 12 more.
* Soft batch, FP32 bulk: a case checkpointed inside a three-case batch
  resumes in a batch of its own and matches the case left in the mix.
* Uneven batch (three grids, FP32): 13 steps, checkpoint, 9 more; the set-up
  layout searches are pinned (they time their candidates, and two steppers
  may pick different layouts, whose round-off differs).
* A state of the wrong shape or type is refused before anything is written.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rigid():
    from pyaxisymflow_amd.stepper import RigidBodyStepper

    nz, nr, z_len = 256, 128, 2.0
    dx = z_len / nz
    z = (np.arange(nz) + 0.5) * dx
    r = (np.arange(nr) + 0.5) * dx
    Z, R = np.meshgrid(z, r)
    chi = (0.1 - np.sqrt((Z - 0.5) ** 2 + R**2) > 0).astype(np.float64)
    return RigidBodyStepper(nz, nr, z_len, 1e-3, 1e4, chi, U_0=1.0, T_ramp=0.05, r_sph=0.1)


def test_rigid_resume_is_bit_identical(tmp_path):
    a = _rigid()
    a.step(31)
    path = tmp_path / "rigid.npz"
    a.save(path)
    b = _rigid()
    b.load(path)
    a.step(20)
    b.step(20)
    assert a.steps_done() == b.steps_done() == 51
    assert np.array_equal(a.vorticity(), b.vorticity())
    assert np.array_equal(a.psi(), b.psi())
    assert np.array_equal(a.scalars(), b.scalars())
    assert np.array_equal(a.drag_history(), b.drag_history())


def test_uneven_batch_resume_is_bit_identical(monkeypatch):
    from pyaxisymflow_amd import workloads
    from pyaxisymflow_amd.stepper import UnevenBatchStepper

    monkeypatch.setenv("AXS_TRIDIAG_LK", "8")
    monkeypatch.setenv("AXS_VMAX_TR", "16")
    cases = []
    for seed, (nz, nr) in enumerate([(128, 64), (192, 96), (96, 48)]):
        sp = workloads.case_spec(seed)
        sp.update(nz=nz, nr=nr)
        _, chi = workloads.case_inputs(sp, nz, nr)
        args = workloads.rigid_case_args(sp, chi)
        args.update(nz=nz, nr=nr, z_len=1.0, r_len=0.5)
        cases.append(args)
    a = UnevenBatchStepper(cases, precision="f32")
    a.step(13)
    snap = a.state()
    b = UnevenBatchStepper(cases, precision="f32")
    b.load_state(snap)
    a.step(9)
    b.step(9)
    for q in range(len(cases)):
        assert a.steps_done(q) == b.steps_done(q) == 22
        assert np.array_equal(a.vorticity(q), b.vorticity(q))
        assert np.array_equal(a.psi(q), b.psi(q))
        assert np.array_equal(a.drag_history(q), b.drag_history(q))


def test_soft_resume_is_bit_identical():
    from pyaxisymflow_amd.stepper import SoftSphereStepper

    a = SoftSphereStepper(grid_size_z=128, G=2.0)
    a.step(17)
    snap = a.state()
    b = SoftSphereStepper(grid_size_z=128, G=2.0)
    b.load_state(snap)
    a.step(12)
    b.step(12)
    assert a.steps_done() == b.steps_done() == 29
    for x, y in zip(a.state().values(), b.state().values()):
        assert np.array_equal(x, y)


def test_soft_batch_case_resumes_in_another_batch():
    from pyaxisymflow_amd.stepper import SoftSphereBatch

    mix = SoftSphereBatch(max_cases=3, precision="f32")
    mix.add(grid_size_z=96, G=0.5)
    case = mix.add(grid_size_z=64, G=3.0)
    mix.add(grid_size_z=80, G=8.0)
    mix.step(9)
    snap = case.state()
    assert snap["vorticity"].dtype == np.float32 and snap["eta1"].dtype == np.float64
    alone = SoftSphereBatch(max_cases=1, precision="f32")
    moved = alone.add(grid_size_z=64, G=3.0)
    moved.load_state(snap)
    mix.step(7)
    alone.step(7)
    assert case.steps_done() == moved.steps_done() == 16
    for k, v in case.state().items():
        assert np.array_equal(v, moved.state()[k]), k


def test_wrong_state_is_refused():
    from pyaxisymflow_amd.stepper import SoftSphereStepper

    a = SoftSphereStepper(grid_size_z=64)
    a.step(2)
    before = a.state()
    bad = dict(before)
    bad["eta1"] = bad["eta1"][:, :-1]
    with pytest.raises(ValueError):
        a.load_state(bad)
    bad = dict(before)
    bad["vorticity"] = bad["vorticity"].astype(np.float32)
    with pytest.raises(ValueError):
        a.load_state(bad)
    bad = dict(before)
    del bad["phi"]
    with pytest.raises(KeyError):
        a.load_state(bad)
    for k, v in a.state().items():
        assert np.array_equal(v, before[k]), k
