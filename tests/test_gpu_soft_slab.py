"""Periodic soft slab on the GPU (PeriodicSoftSlabStepper,
examples/PeriodicSoftSlab/periodic_soft_slab.py:43-404: an elastic layer
around the axis, its reference map extrapolated by the static PDE scheme of
bounded_static_PDE_extrapolation.py, driven by an oscillating annular wall;
z-periodic with ghost columns).

* Against the oracle's restatement of the same loop
  (oracle/drivers.py PeriodicSoftSlabCase, exact Poisson route): 100 steps,
  vorticity, psi, u_z and both reference-map components within 1e-10
  relative L2 (the transforms and the radial solve differ from the oracle's
  in round-off; the wall's cos(omega t) is the device's, within an ulp of
  numpy's), u_r against u_z's size, the same times, and the same Jacobi
  iteration counts in every extrapolation (the residual's summation order
  differs from numpy's; the counts must still agree).
* Graph replay equals eager stepping bit for bit.
* Physics (the driver's own check, theory_soft_slab.py:15-87 against the
  samples of :229-238): over the third period the axial velocity for
  r < R_tube in column nz / 2 follows the Kelvin-Voigt theory within 3% of
  V_wall at every sample (measured: at most 1.5% from the second period on,
  1.4% by the fifteenth); the same comparison a quarter period off gives at
  least 40%, so a phase, sign or amplitude mistake fails the bar.
Grid 256 x 8 (the driver's default), Re = 10, Er = 0.25.
"""
import numpy as np
import pytest

from oracle.drivers import PeriodicSoftSlabCase, soft_slab_theory

pytestmark = pytest.mark.gpu


def rel_l2(a, b, scale=None):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b if scale is None else scale), 1e-300)


def test_soft_slab_100_steps_match_oracle():
    from pyaxisymflow_amd.stepper import PeriodicSoftSlabStepper

    print("soft slab: 100 steps vs oracle (exact route), 256 x 8", flush=True)
    cpu = PeriodicSoftSlabCase(solver="exact")
    gpu = PeriodicSoftSlabStepper()
    for k in range(100):
        cpu.step()
        gpu.step(1)
        if k in (0, 1, 50, 99):
            assert gpu.jacobi_last() == sum(cpu.extrapolator.iterations[-4:]), k
    uz, ur = gpu.velocity()
    e1, e2 = gpu.eta()
    errs = dict(w=rel_l2(gpu.vorticity(), cpu.vorticity), psi=rel_l2(gpu.psi(), cpu.psi),
                uz=rel_l2(uz, cpu.u_z), ur=rel_l2(ur, cpu.u_r, cpu.u_z),
                eta1=rel_l2(e1, cpu.eta1), eta2=rel_l2(e2, cpu.eta2))
    print("soft slab 100 steps: " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert gpu.steps_done() == 100
    assert abs(gpu.time() - cpu.t) <= 1e-13
    assert gpu.jacobi_capped() == 0
    for k, v in errs.items():
        assert v < 1e-10, k
    ts, prof = gpu.samples()
    assert len(ts) == len(cpu.samples) == 1 and ts[0] == 0.0


def test_soft_slab_graph_equals_eager():
    from pyaxisymflow_amd.stepper import PeriodicSoftSlabStepper

    a = PeriodicSoftSlabStepper(use_graph=True)
    b = PeriodicSoftSlabStepper(use_graph=False)
    a.step(37)  # 16-step graphs plus single-step remainders
    b.step(37)
    assert np.array_equal(a.vorticity(), b.vorticity())
    assert np.array_equal(a.psi(), b.psi())
    for x, y in zip(a.eta(), b.eta()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.scalars(), b.scalars())


def test_soft_slab_follows_theory():
    from pyaxisymflow_amd.stepper import PeriodicSoftSlabStepper

    st = PeriodicSoftSlabStepper(t_end=3.0)
    while st.time() < 3.0 - 1e-12:
        st.step(4096)
        print(f"soft slab t = {st.time():.3f} ({st.steps_done()} steps)", flush=True)
    ts, prof = st.samples()
    theory, _, _ = soft_slab_theory(st.L_f, st.L_s, st.shear_rate, st.omega, st.G, st.V_wall,
                                    st.rho_f, st.nu, resolution=st.r[:st.sample_rows])
    sel = ts >= 2.0 - 1e-9
    assert sel.sum() >= 9
    err = max(np.max(np.abs(p - theory(t))) for t, p in zip(ts[sel], prof[sel])) / st.V_wall
    off = min(np.max(np.abs(p - theory(t + 0.25))) for t, p in zip(ts[sel], prof[sel])) / st.V_wall
    print(f"soft slab third period: max |u - theory| / V = {err:.4f}, quarter period off "
          f"{off:.3f}", flush=True)
    assert err < 0.03
    assert off > 0.40
    assert st.jacobi_capped() == 0
