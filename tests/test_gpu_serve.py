"""Serving form (continuous batching of soft-sphere requests).

* A request's results are the same alone and in the mix (FP32 bulk, the
  server's default, and FP64): the admitted
  requests are cases of one soft batch (one launch per pass over all of
  them) whose arithmetic per case does not depend on the others, so the
  final fields of a request run beside others, joining and leaving at
  different chunks, equal, bit for bit, those of the same request run alone.
* Every request completes its steps; time-to-solution >= its stepping; the
  slot and memory limits hold (admission waits instead of overcommitting).
* A request that fails (a NaN planted in its vorticity at admission: its
  first step's velocity maximum, and so its dt, is NaN, as numpy's amax
  makes it in the reference's dt rule) is evicted after its first chunk and
  recorded as failed; the requests beside it complete with the bits of
  their solo runs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _reqs():
    base = dict(nu=1e-3, rho_f=1.0, r_ball=0.15, steps=40)
    return [dict(id=0, arrival=0.0, grid_size_z=128, nr=64, G=1.0, **base),
            dict(id=1, arrival=0.0, grid_size_z=96, nr=48, G=5.0, **base),
            dict(id=2, arrival=0.001, grid_size_z=160, nr=80, G=0.3, **base),
            dict(id=3, arrival=0.002, grid_size_z=128, nr=64, G=8.0, **base)]


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_request_same_alone_and_in_the_mix(precision):
    from pyaxisymflow_amd.serve import SoftServer

    reqs = _reqs()
    alone = SoftServer(max_active=1, chunk=7, precision=precision).run([reqs[0]],
                                                                        keep_state=True)[0]
    mix = SoftServer(max_active=4, chunk=5, precision=precision).run(reqs, keep_state=True)
    got = [r for r in mix if r["id"] == 0][0]
    assert got["steps"] == alone["steps"] == 40
    for k in ("vorticity", "psi", "eta1", "eta2", "phi", "scalars"):
        assert np.array_equal(got["state"][k], alone["state"][k]), k


def test_serving_limits_and_records():
    from pyaxisymflow_amd.serve import SoftServer, footprint_bytes, latency_summary

    reqs = _reqs()
    budget = footprint_bytes(reqs[2], "f32") + footprint_bytes(reqs[0], "f32")  # two at once
    recs = SoftServer(max_active=3, hbm_budget=budget, chunk=10).run(reqs)
    assert [r["id"] for r in recs] == [0, 1, 2, 3]
    for r in recs:
        assert r["steps"] == 40 and r["sim_time"] > 0
        assert r["finished"] >= r["admitted"] >= r["arrival"] - 1e-9
        assert r["time_to_solution"] == pytest.approx(r["finished"] - r["arrival"])
    s = latency_summary(recs)
    assert s["requests"] == 4 and s["p99_s"] >= s["p50_s"] > 0


def test_request_chunks_do_not_wait_on_each_other():
    """Single-case steppers stepped with join=False queue on their own
    streams only: a short chunk of one completes while a long chunk of
    another, queued before it, is still running."""
    import torch

    from pyaxisymflow_amd.stepper import SoftSphereStepper

    long_run = SoftSphereStepper(grid_size_z=256, G=1.0)
    short_run = SoftSphereStepper(grid_size_z=128, G=1.0)
    # capture both steppers' graphs before timing anything
    long_run.step(1, join=False)
    short_run.step(1, join=False)
    torch.cuda.synchronize()
    long_run.step(400, join=False)
    ev_long = torch.cuda.Event()
    ev_long.record(long_run.stream)
    short_run.step(2, join=False)
    ev_short = torch.cuda.Event()
    ev_short.record(short_run.stream)
    ev_short.synchronize()
    assert not ev_long.query(), "the short chunk waited for the long one"
    ev_long.synchronize()
    assert long_run.steps_done() == 401 and short_run.steps_done() == 3


def test_failed_request_is_evicted_and_the_rest_unaffected():
    from pyaxisymflow_amd.serve import SoftServer, latency_summary

    reqs = _reqs()

    def plant(r, case):
        if r["id"] == 1:
            case.t_vorticity[case.nr // 2, case.nz // 2] = float("nan")

    print("serving: NaN planted in request 1 of 4", flush=True)
    alone = {q["id"]: SoftServer(max_active=1, chunk=7).run([q], keep_state=True)[0]
             for q in reqs if q["id"] != 1}
    mix = {r["id"]: r for r in SoftServer(max_active=4, chunk=5).run(reqs, keep_state=True,
                                                                      on_admit=plant)}
    bad = mix[1]
    assert bad["failed"] and bad["steps"] == 5 and not np.isfinite(bad["sim_time"])
    for i, ref in alone.items():
        got = mix[i]
        assert not got["failed"] and got["steps"] == 40
        for k in ("vorticity", "psi", "eta1", "eta2", "phi", "scalars"):
            assert np.array_equal(got["state"][k], ref["state"][k]), (i, k)
    s = latency_summary(list(mix.values()))
    assert s["requests"] == 3 and s["failed"] == 1
