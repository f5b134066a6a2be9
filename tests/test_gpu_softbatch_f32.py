"""FP32-bulk soft batch (axs_soft_batch_create_f32 / axs_soft_batch_add_f32):
the serving form's precision.  The flow fields (vorticity, psi, velocity,
the solid mask) and their passes are FP32; the radial elimination, the
reference map, the level set, its reinit and the least-squares
extrapolation stay FP64.

* Against the FP64 oracle (oracle/drivers.py SoftSphereCase, exact Poisson
  route) on the Taylor-Green soft-sphere loop
  (soft_sphere_in_taylor_green_vortex.py:101-213) at nz = 128, the case run
  among two others.  The loop amplifies round-off (the level set's and the
  extrapolation's thresholds move cells between zones), so that after 200
  steps the FP64 batch is itself ~1.2e-4 from the oracle in omega and the
  oracle's two Poisson routes ~0.9e-4 from each other.  Bars: after 50 steps
  omega, psi, eta and phi within 1e-4 relative L2 (the FP32 bar; measured
  at most 6.8e-6) and the simulated time within 1e-6 (dt follows max|u|,
  which is FP32); after 200 steps every field within 1e-4, or, where the
  amplification has carried the FP64 batch itself past that, within twice
  the FP64 batch's own distance from the oracle (measured: omega 1.30e-4
  against 1.18e-4).  The solid mask's support (chi > 0) equals the
  oracle's at both checks.
* A case gives the same bits alone and in a mix (the per-case arithmetic is
  fixed), also when cases join and leave; graph replay equals eager steps.
* An FP32 batch refuses FP64 buffers and an FP64 batch FP32 ones.
"""
import ctypes

import numpy as np
import pytest

from oracle.drivers import SoftSphereCase

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _fields(c):
    if isinstance(c, SoftSphereCase):
        return dict(w=c.vorticity, psi=c.psi, e1=c.eta1, e2=c.eta2, phi=c.ball_phi)
    e1, e2 = c.eta()
    return dict(w=c.vorticity(), psi=c.psi(), e1=e1, e2=e2, phi=c.phi())


def _assert_same_bits(a, b, what):
    fa, fb = _fields(a), _fields(b)
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), f"{what}: {k}"
    assert np.array_equal(a.char_func(), b.char_func()), what
    assert np.array_equal(a.scalars()[:9], b.scalars()[:9]), what


def test_f32_batch_against_fp64_oracle():
    from pyaxisymflow_amd.stepper import SoftSphereBatch

    print("soft batch f32: 50 + 150 steps vs FP64 oracle, nz 128 among 96 and 160", flush=True)
    cpu = SoftSphereCase(grid_size_z=128, solver="exact")
    eig = SoftSphereCase(grid_size_z=128, solver="eig")
    b32 = SoftSphereBatch(max_cases=3, precision="f32")
    b32.add(grid_size_z=96, G=0.5)
    c32 = b32.add(grid_size_z=128)
    b32.add(grid_size_z=160, G=5.0)
    b64 = SoftSphereBatch(max_cases=1)
    c64 = b64.add(grid_size_z=128)
    assert str(c32.t_vorticity.dtype) == str(c32.t_char_func.dtype) == "torch.float32"
    assert str(c32.t_psi.dtype) == "torch.float32"
    assert str(c32.t_eta1.dtype) == str(c32.t_phi.dtype) == "torch.float64"
    done = 0
    for steps in (50, 200):
        for _ in range(steps - done):
            cpu.step()
            eig.step()
        b32.step(steps - done)
        b64.step(steps - done)
        done = steps
        g32, g64, ref, alt = _fields(c32), _fields(c64), _fields(cpu), _fields(eig)
        errs, errs64 = {}, {}
        for k in ref:
            errs[k] = rel_l2(g32[k].astype(np.float64), ref[k])
            errs64[k] = rel_l2(g64[k], ref[k])
            print(f"soft batch {steps} steps {k}: f32 {errs[k]:.2e}  f64 {errs64[k]:.2e}  "
                  f"oracle routes {rel_l2(alt[k], ref[k]):.2e}", flush=True)
        flips32 = int(np.sum((c32.char_func() > 0) != (cpu.ball_char_func > 0)))
        flips64 = int(np.sum((c64.char_func() > 0) != (cpu.ball_char_func > 0)))
        print(f"soft batch {steps} steps: mask support differs in {flips32} cells (f32), "
              f"{flips64} (f64)", flush=True)
        assert c32.steps_done() == steps and c32.lsq_unconverged() == 0
        assert flips32 == 0
        if steps == 50:
            assert abs(c32.time() - cpu.t) <= 1e-6 * cpu.t
            for k, v in errs.items():
                assert v < 1e-4, k
        else:
            for k, v in errs.items():
                assert v < 1e-4 or v <= 2 * errs64[k], k


def test_f32_case_alone_equals_in_mix_and_join_leave():
    from pyaxisymflow_amd.stepper import SoftSphereBatch

    print("soft batch f32: alone vs mix with join/leave, nz 64, 5 + 9 steps", flush=True)
    alone = SoftSphereBatch(max_cases=1, precision="f32")
    a = alone.add(grid_size_z=64, G=2.0)
    alone.step(14)
    mix = SoftSphereBatch(max_cases=3, precision="f32")
    other = mix.add(grid_size_z=100, G=0.3)
    m = mix.add(grid_size_z=64, G=2.0)
    mix.step(5)
    mix.remove(other)
    late = mix.add(grid_size_z=48, G=8.0)
    mix.step(9)
    _assert_same_bits(a, m, "alone vs mix")
    assert m.steps_done() == 14 and late.steps_done() == 9


def test_f32_graph_equals_eager():
    from pyaxisymflow_amd.stepper import SoftSphereBatch

    print("soft batch f32: graph vs eager, nz 64 + 96, 19 steps", flush=True)
    g = SoftSphereBatch(max_cases=2, use_graph=True, precision="f32")
    e = SoftSphereBatch(max_cases=2, use_graph=False, precision="f32")
    ga, gb = g.add(grid_size_z=64), g.add(grid_size_z=96, G=0.5)
    ea, eb = e.add(grid_size_z=64), e.add(grid_size_z=96, G=0.5)
    g.step(19)
    e.step(19)
    _assert_same_bits(ga, ea, "graph vs eager (64)")
    _assert_same_bits(gb, eb, "graph vs eager (96)")


def test_batches_refuse_the_other_precision():
    from pyaxisymflow_amd import _lib
    from pyaxisymflow_amd.stepper import SoftSphereBatch, _SoftCaseView, _soft_case_state

    lib = _lib.load()
    b32 = SoftSphereBatch(max_cases=1, precision="f32")
    b64 = SoftSphereBatch(max_cases=1)
    v64, v32 = _SoftCaseView(b64), _SoftCaseView(b32)
    _soft_case_state(v64, 64, 1.0, 1e-3, 1.0, 0.15, 0.5, 5e-2, 3)
    _soft_case_state(v32, 64, 1.0, 1e-3, 1.0, 0.15, 0.5, 5e-2, 3, precision="f32")
    cid = ctypes.c_int()
    st = lib.axs_soft_batch_add(b32._h, ctypes.byref(v64.cfg), ctypes.byref(v64._bufs),
                                ctypes.byref(cid))
    assert st == _lib.AXS_ERR_INVALID and b"add_f32" in lib.axs_last_error()
    st = lib.axs_soft_batch_add_f32(b64._h, ctypes.byref(v32.cfg), ctypes.byref(v32._bufs),
                                    ctypes.byref(cid))
    assert st == _lib.AXS_ERR_INVALID and b"axs_soft_batch_add" in lib.axs_last_error()
    assert b32.info()[0] == 0 and b64.info()[0] == 0
