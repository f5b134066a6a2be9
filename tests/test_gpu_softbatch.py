"""Soft batch (axs_soft_batch_*): many soft-sphere runs
(soft_sphere_in_taylor_green_vortex.py:101-213) stepped together, one
launch per pass over all of them.

* One step of a case in a mixed batch against the oracle's exact route
  (oracle/drivers.py SoftSphereCase): every field within 1e-12 relative L2,
  dt within 1e-12 (the single-case stepper's bar).
* Twelve steps (from the second step on, the passes after the mirror visit
  only the extrapolation zone's rows and a margin): every field no further
  from the oracle's exact route than twice the oracle's own exact-vs-eig
  spread (the single-case stepper's bar; the loop amplifies round-off).
* A case gives the same bits alone and in a mix of other grids and solids,
  also when cases join and leave between runs and beside a long-row case
  that widens the row workgroups, and graph replay equals eager stepping.
* Long rows: one step of cases whose row FFT has a stage of 300 butterflies
  (nz 580: the chirp-z length 600; nz 1200: the direct length 600, both
  with a radix-2 stage), in FP64 and FP32 against the oracle: the batch's
  row workgroup covers every butterfly of the widest stage (a butterfly
  left out puts an O(1) error into psi).  Bars: FP64 1e-10 relative L2,
  below the oracle's own exact-vs-eig spread after one step at these grids
  (omega 6.0e-10 at nz 580, 2.6e-10 at nz 1200; measured on the GPU:
  omega 8.8e-12, psi 9.8e-14 at nz 580); FP32 1e-4.
Grids 48 x 24 to 160 x 80 (nz = 2 nr, as the extrapolation requires), and
580 x 290 and 1200 x 600 for the long rows.
"""
import numpy as np
import pytest

from oracle.drivers import SoftSphereCase

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _fields(c):
    if isinstance(c, SoftSphereCase):
        return dict(w=c.vorticity, psi=c.psi, e1=c.eta1, e2=c.eta2, phi=c.ball_phi,
                    chi=c.ball_char_func)
    e1, e2 = c.eta()
    return dict(w=c.vorticity(), psi=c.psi(), e1=e1, e2=e2, phi=c.phi(), chi=c.char_func())


def _assert_same_bits(a, b, what):
    fa, fb = _fields(a), _fields(b)
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), f"{what}: {k}"
    assert np.array_equal(a.scalars()[:9], b.scalars()[:9]), what


def test_batch_case_one_step_matches_oracle():
    from pyaxisymflow_amd.stepper import SoftSphereBatch

    print("soft batch: 1 step vs oracle, nz 128 among 96 and 160", flush=True)
    cpu = SoftSphereCase(grid_size_z=128, solver="exact")
    b = SoftSphereBatch(max_cases=4)
    b.add(grid_size_z=96, G=0.5)
    case = b.add(grid_size_z=128)
    b.add(grid_size_z=160, G=5.0)
    cpu.step()
    b.step(1)
    g, c = _fields(case), _fields(cpu)
    errs = {k: rel_l2(g[k], c[k]) for k in c}
    print("soft batch 1 step: " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    sc = case.scalars()
    assert case.steps_done() == 1
    assert abs(sc[1] - cpu.dts[-1]) <= 1e-12 * cpu.dts[-1]
    assert case.lsq_unconverged() == 0
    for k, v in errs.items():
        assert v < 1e-12, k


@pytest.mark.parametrize("precision,bar", [("f64", 1e-10), ("f32", 1e-4)])
def test_batch_long_rows_one_step_match_oracle(precision, bar):
    from pyaxisymflow_amd.stepper import SoftSphereBatch

    print(f"soft batch {precision}: 1 step vs oracle, nz 580 and 1200 among 96", flush=True)
    b = SoftSphereBatch(max_cases=3, precision=precision)
    b.add(grid_size_z=96, G=0.5)
    cases = {nz: b.add(grid_size_z=nz, G=2.0) for nz in (580, 1200)}
    b.step(1)
    for nz, case in cases.items():
        cpu = SoftSphereCase(grid_size_z=nz, G=2.0, solver="exact")
        cpu.step()
        g, c = _fields(case), _fields(cpu)
        errs = {k: rel_l2(np.asarray(g[k], np.float64), c[k]) for k in c}
        print(f"soft batch {precision} nz {nz}: " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()),
              flush=True)
        assert case.lsq_unconverged() == 0
        for k, v in errs.items():
            assert v < bar, (nz, k)


def test_batch_case_12_steps_within_oracle_route_spread():
    from pyaxisymflow_amd.stepper import SoftSphereBatch

    steps = 12
    print("soft batch: 12 steps vs oracle routes, nz 128 among 96 and 160", flush=True)
    exact = SoftSphereCase(grid_size_z=128, solver="exact")
    eig = SoftSphereCase(grid_size_z=128, solver="eig")
    b = SoftSphereBatch(max_cases=3)
    b.add(grid_size_z=96, G=0.5)
    case = b.add(grid_size_z=128)
    b.add(grid_size_z=160, G=5.0)
    for _ in range(steps):
        exact.step()
        eig.step()
    b.step(steps)
    g, c, e = _fields(case), _fields(exact), _fields(eig)
    for k in c:
        got, spread = rel_l2(g[k], c[k]), rel_l2(e[k], c[k])
        print(f"soft batch {steps} steps {k}: gpu {got:.2e} oracle routes {spread:.2e}", flush=True)
        assert got <= 2 * spread + 1e-14, k
    assert abs(case.time() - exact.t) <= 1e-12 * exact.t
    assert case.lsq_unconverged() == 0


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_small_case_same_bits_beside_a_long_row_case(precision):
    """A long-row case (nz 1200: a row stage of 300 butterflies) widens the
    batch's row workgroups from 64 to 128 threads; a small case's rows give
    the same bits either way (each butterfly's arithmetic does not depend
    on which thread computes it)."""
    from pyaxisymflow_amd.stepper import SoftSphereBatch

    print(f"soft batch {precision}: nz 96 alone vs beside nz 1200, 6 steps", flush=True)
    alone = SoftSphereBatch(max_cases=1, precision=precision)
    a = alone.add(grid_size_z=96, G=0.5)
    mix = SoftSphereBatch(max_cases=2, precision=precision)
    mix.add(grid_size_z=1200, G=2.0)
    m = mix.add(grid_size_z=96, G=0.5)
    alone.step(6)
    mix.step(6)
    _assert_same_bits(a, m, "alone vs beside a long-row case")


def test_batch_case_alone_equals_in_mix():
    from pyaxisymflow_amd.stepper import SoftSphereBatch

    print("soft batch: alone vs mix, nz 64, 12 steps", flush=True)
    alone = SoftSphereBatch(max_cases=1)
    a = alone.add(grid_size_z=64, G=2.0)
    mix = SoftSphereBatch(max_cases=4)
    mix.add(grid_size_z=100, G=0.3)
    m = mix.add(grid_size_z=64, G=2.0)
    mix.add(grid_size_z=48, G=8.0)
    alone.step(12)
    mix.step(12)
    _assert_same_bits(a, m, "alone vs mix")
    assert a.steps_done() == 12


def test_batch_join_leave_between_runs():
    from pyaxisymflow_amd.stepper import SoftSphereBatch

    print("soft batch: join/leave, nz 64, 5 + 9 steps", flush=True)
    alone = SoftSphereBatch(max_cases=1)
    a = alone.add(grid_size_z=64)
    alone.step(14)
    b = SoftSphereBatch(max_cases=3)
    other = b.add(grid_size_z=80, G=0.2)
    x = b.add(grid_size_z=64)
    b.step(5)
    b.remove(other)
    late = b.add(grid_size_z=120, G=3.0)
    b.step(9)
    _assert_same_bits(a, x, "join/leave")
    assert late.steps_done() == 9 and x.steps_done() == 14
    n, queued, launches, captures = b.info()
    print(f"soft batch info: cases {n} queued sweeps {queued} graph launches {launches} "
          f"captures {captures}", flush=True)
    assert n == 2 and captures >= 2


def test_batch_graph_equals_eager():
    from pyaxisymflow_amd.stepper import SoftSphereBatch

    print("soft batch: graph vs eager, nz 64 + 96, 19 steps", flush=True)
    g = SoftSphereBatch(max_cases=2, use_graph=True)
    e = SoftSphereBatch(max_cases=2, use_graph=False)
    ga, gb = g.add(grid_size_z=64), g.add(grid_size_z=96, G=0.5)
    ea, eb = e.add(grid_size_z=64), e.add(grid_size_z=96, G=0.5)
    g.step(19)  # eight-step graphs plus single-step remainders
    e.step(19)
    _assert_same_bits(ga, ea, "graph vs eager (64)")
    _assert_same_bits(gb, eb, "graph vs eager (96)")
