"""Serving form: soft-sphere simulations as requests, continuously batched
onto one GPU.

A request is one run of the soft-sphere loop
(examples/SoftSphereInTaylorGreenVortex/soft_sphere_in_taylor_green_vortex.py:101-213)
for a fixed number of steps, with its own grid and shear modulus.  The
server admits queued requests while a slot is free and their device
footprint fits the HBM budget, as cases of one SoftSphereBatch: every pass
of a step is one launch over all admitted requests.  The batch runs in
chunks of graph-replayed steps; between chunks finished requests leave and
arrivals join (continuous batching).  A chunk is as long as the shortest
remaining run (at most `chunk` steps), so no request takes a step beyond its
own count.  A case's arithmetic does not depend on the other cases, so a
request's results are the same bits alone and in any mix.

precision "f32" (the default) serves with the FP32-bulk soft batch: the
flow fields and their passes in FP32, the radial elimination, the
reference map, the level set and the extrapolation in FP64; "f64" runs
every pass in FP64, as the reference's loop does.

Time-to-solution of a request = completion time - arrival time (queueing,
set-up, and stepping), measured on the host clock at the end of the chunk
that completes it.

Failure handling: after every chunk the server reads each active request's
time and dt; a request whose time or dt is no longer finite and positive
(a NaN or Inf velocity makes dt NaN, as in the reference's dt rule) is
evicted at once and recorded as failed.  The other requests are unaffected:
no pass mixes values of different cases.
"""
import time

import numpy as np

from pyaxisymflow_amd import _lib


def request_set(n, seed=0, rate=1.0, steps=2000, nr_range=(128, 1024)):
    """Synthetic requests: Poisson arrivals at `rate` per second; nr =
    clip(round(lognormal(ln 256, 0.5)), 128, 1024), nz = 2 nr (any size: the
    row transforms take every nz); G ~ U[0.1, 10]; rho_f = 1, nu = 1e-3,
    r_ball = 0.15, the driver's Taylor-Green vorticity."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n)
    arrivals = np.cumsum(gaps) - gaps[0]
    out = []
    for q in range(n):
        nr = int(np.clip(round(rng.lognormal(np.log(256), 0.5)), *nr_range))
        out.append(dict(id=q, arrival=float(arrivals[q]), grid_size_z=2 * nr, nr=nr,
                        G=float(rng.uniform(0.1, 10.0)), nu=1e-3, rho_f=1.0, r_ball=0.15,
                        steps=int(steps)))
    return out


def footprint_bytes(req, precision="f64"):
    """Device bytes of one soft request: 27 (nr, nz) fields (caller-owned
    and scratch; in "f32" the 18 flow fields -- vorticity, psi, u_z, u_r,
    char_func and 13 scratch -- are float32, the rest float64), the
    doubled grid (eta x2, LS fits x2: FP64; flags x2: int16; pending: int8)
    and the spectral tables."""
    n = req["nr"] * req["grid_size_z"]
    fields = 4 * n * 18 + 8 * n * 9 if precision == "f32" else 8 * n * (9 + 18)
    return fields + 2 * n * (4 * 8 + 2 * 2 + 1) + 8 * n


class SoftServer:
    """Continuous batching of soft-sphere requests on one GPU.

    max_active: slots (concurrent requests); hbm_budget: bytes the active
    requests may hold (default: 80% of free device memory); chunk: most
    steps run between scheduling decisions; precision: "f32" (FP32 bulk) or
    "f64"."""

    def __init__(self, max_active=64, hbm_budget=None, chunk=100, precision="f32"):
        torch = _lib.torch()
        dev = _lib.device()
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(dev)
            hbm_budget = int(0.8 * free)
        self.max_active = int(max_active)
        self.hbm_budget = int(hbm_budget)
        self.chunk = int(chunk)
        if precision not in ("f32", "f64"):
            raise ValueError("precision must be 'f32' or 'f64'")
        self.precision = precision

    def run(self, requests, keep_state=False, on_admit=None):
        """Serves `requests` (dicts from request_set) with arrivals relative
        to the start of the call; returns one record per request (arrival,
        admitted, finished, time_to_solution, steps, final time, failed)
        and, with keep_state, each request's final fields.  on_admit(request,
        case), if given, is called once a request has joined the batch and
        before its first step (to set a custom initial state in the case's
        tensors)."""
        torch = _lib.torch()
        from pyaxisymflow_amd.stepper import SoftSphereBatch

        pending = sorted(requests, key=lambda r: r["arrival"])

        def fp(r):
            return footprint_bytes(r, self.precision)

        for r in pending:
            if fp(r) > self.hbm_budget:
                raise ValueError(f"request {r['id']} needs {fp(r)} bytes, more "
                                 f"than the budget {self.hbm_budget}")
        batch = SoftSphereBatch(max_cases=self.max_active, precision=self.precision)
        self.last_batch_info = None
        active, records, used = [], {}, 0
        t0 = time.perf_counter()
        while pending or active:
            now = time.perf_counter() - t0
            # admission: arrived, slot free, memory fits (first come first served)
            while (pending and pending[0]["arrival"] <= now and len(active) < self.max_active
                   and used + fp(pending[0]) <= self.hbm_budget):
                r = pending.pop(0)
                case = batch.add(grid_size_z=r["grid_size_z"], G=r["G"], nu=r["nu"],
                                 rho_f=r["rho_f"], r_ball=r["r_ball"])
                if on_admit is not None:
                    on_admit(r, case)
                    torch.cuda.synchronize()  # its writes land before the first step
                used += fp(r)
                active.append(dict(req=r, case=case, done=0,
                                   admitted=time.perf_counter() - t0))
            if not active:
                wait = pending[0]["arrival"] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.01))
                continue
            k = min(self.chunk, min(a["req"]["steps"] - a["done"] for a in active))
            batch.step(k, join=False)
            ev = torch.cuda.Event()
            ev.record(batch.stream)
            ev.synchronize()
            finished = time.perf_counter() - t0
            # time and dt of every active request in one read
            health = torch.stack([a["case"].t_scalars[:2] for a in active]).cpu().numpy()
            for a, (t_now, dt_now) in zip(list(active), health):
                a["done"] += k
                r = a["req"]
                failed = not (np.isfinite(t_now) and np.isfinite(dt_now) and dt_now > 0)
                if a["done"] < r["steps"] and not failed:
                    continue
                c = a["case"]
                rec = dict(id=r["id"], nr=r["nr"], arrival=r["arrival"], admitted=a["admitted"],
                           finished=finished, steps=a["done"], sim_time=float(t_now),
                           lsq_unconverged=c.lsq_unconverged(), failed=failed)
                rec["time_to_solution"] = rec["finished"] - rec["arrival"]
                if keep_state:
                    e1, e2 = c.eta()
                    rec["state"] = dict(vorticity=c.vorticity(), psi=c.psi(), eta1=e1, eta2=e2,
                                        phi=c.phi(), scalars=c.scalars())
                records[r["id"]] = rec
                used -= fp(r)
                batch.remove(c)
                active.remove(a)
        self.last_batch_info = batch.info()
        del batch
        return [records[k] for k in sorted(records)]


def latency_summary(records):
    """Percentiles over the requests that completed (failed ones counted
    apart)."""
    done = [r for r in records if not r.get("failed")]
    tts = np.array([r["time_to_solution"] for r in done])
    return dict(requests=len(done), failed=len(records) - len(done),
                p50_s=float(np.percentile(tts, 50)), p99_s=float(np.percentile(tts, 99)),
                mean_s=float(tts.mean()), max_s=float(tts.max()))
