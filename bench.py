"""Benchmark: Mcell-steps/s of the plain FP64 timestep on one MI355X.

Workload (kernel-scale plain FP64, nz x nr = 2048 x 1024): z in [0, 2],
dx = 2/nz, r in [0, 1]; vorticity = 64 seeded Gaussians (amplitudes N(0,1),
widths U[0.01, 0.05], centres uniform); body mask = smooth Heaviside of the
union of 8 on-axis spheres (radius U[0.02, 0.06], centre z U[0.2, 1.8]);
lambda = 1e4, nu = 1.5e-3, free stream U = 1 (no ramp).  A step is the full
kill -> Poisson -> velocity -> dt -> penalise -> drag -> ENO3 -> RK2 sequence,
device resident and replayed from a hipGraph.

Prints one JSON line (rank 0).  `value` is whole-job Mcell-steps/s with the
fields resident in HBM.  `roofline` covers the dominant stage, timed live
with HIP events on the stepper's stream; `cpu_baseline` times the CPU oracle
(numpy restatement of the reference, dense eig-route Poisson) on a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-steps/s @2048x1024 FP64 vs batched FP32; p50/p99 case time-to-solution"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per cell of each stage (compulsory reads + writes, FP64):
STAGE_BYTES_PER_CELL = {
    "taper_dct2_rows": 16,   # read omega (taper applied on the fly); write omega-hat
    "tridiag_columns": 16,   # read omega-hat, write psi-hat
    "dct3_rows": 16,         # read psi-hat, write psi
    "velocity_max": 8,       # read psi
    "body_mask": 8,          # write chi (moving body only)
    "step_fused": 32,        # read psi, chi, omega; write omega
    # implicit-diffusion drivers only: DST-I rows, radial solve, DST-I rows
    # (16 B each; the DST-I matrix stays in L2 for nz <= 1024)
    "implicit_diffusion": 48,
}


def profile_traffic(stage, nz, nr, dtype="double"):
    """HBM bytes per launch of a stage's kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE and WRITE_SIZE, each calibrated against the
    copy kernel), or None when no summary covers this grid and kernel."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_fp64_{nz}x{nr}.md")
    if not os.path.exists(path):
        return None
    key = f"{stage}_kernel<{dtype}"
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 4 and key in cells[0].replace("`", ""):
            try:
                return round(float(cells[3]) * nz * nr)
            except ValueError:
                return None
    return None


def workload(nz, nr, seed=0):
    from pyaxisymflow_amd.workloads import kernel_scale_workload

    return kernel_scale_workload(nz, nr, seed)


def time_gpu(nz, nr, steps, warmup, stage_reps):
    import ctypes

    import torch

    from pyaxisymflow_amd import _lib
    from pyaxisymflow_amd.stepper import RigidBodyStepper

    wl = workload(nz, nr)
    st = RigidBodyStepper(nz, nr, wl["z_len"], wl["nu"], wl["brink_lam"], wl["char_func"],
                          U_0=wl["U_0"], T_ramp=None, r_sph=wl["r_sph"],
                          vorticity=wl["vorticity"], history=steps + warmup + 8)
    stream = st.stream
    # the stepper's own stream carries every kernel and the device is
    # synchronized on both sides, so the steps need no ordering against
    # torch's current stream (join=False: a cross-stream wait would put
    # ~27 us of idle device time at the start of the timed region)
    st.step(warmup, join=False)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    st.step(steps, join=False)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    w = st.vorticity()
    finite = bool(np.isfinite(w).all())
    cd = st.drag_history()
    # per-stage timing, each stage launched alone on the same stream
    lib = _lib.load()
    stages = {}
    for q in range(lib.axs_rigid_stepper_stage_count()):
        name = lib.axs_rigid_stepper_stage_name(q).decode()
        _lib.call("axs_rigid_stepper_run_stage", st._h, q, 3, ctypes.c_void_p(stream.cuda_stream))
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        _lib.call("axs_rigid_stepper_run_stage", st._h, q, stage_reps,
                  ctypes.c_void_p(stream.cuda_stream))
        b.record(stream)
        torch.cuda.synchronize()
        stages[name] = a.elapsed_time(b) * 1e3 / stage_reps  # us per launch
    # plain copy kernel: achievable-bandwidth reference and byte-counter calibration
    n = 1 << 24  # 128 MiB per buffer
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    args = (ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()), n,
            ctypes.c_void_p(stream.cuda_stream))
    _lib.call("axs_device_copy_f64", *args)
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    reps = 20
    a.record(stream)
    for _ in range(reps):
        _lib.call("axs_device_copy_f64", *args)
    b.record(stream)
    torch.cuda.synchronize()
    copy_gbs = 2 * 8 * n * reps / (a.elapsed_time(b) * 1e-3) / 1e9
    launch = "graph" if st.use_graph else "eager"
    return ms, finite, float(cd[-1]) if len(cd) else None, stages, copy_gbs, launch


def time_batched(steps, warmup, seeds=range(64)):
    """The batched form: the 64-case set of pyaxisymflow_amd.workloads
    (grids nr in {128, 192, 256, 384, 512}, nz = 2 nr; sphere, prolate
    spheroid and torus bodies; Re in [10, 500]) as one UnevenBatchStepper:
    every pass is one launch over all 64 cases, 16 steps per graph launch.
    FP32 is timed; the same cases in FP64 give the drag error.  The grouped
    layout (one batched stepper per grid, each on its own stream) is timed
    beside it."""
    import torch

    from pyaxisymflow_amd import workloads
    from pyaxisymflow_amd.stepper import RigidBatchStepper, UnevenBatchStepper

    groups = workloads.grouped_cases(seeds)
    hist = steps + warmup + 4
    cases = []
    for (nz, nr), specs in groups.items():
        for sp in specs:
            a = workloads.rigid_case_args(sp, workloads.case_inputs(sp, nz, nr)[1])
            a.update(nz=nz, nr=nr, z_len=1.0, r_len=0.5)
            cases.append(a)
    b32 = UnevenBatchStepper(cases, precision="f32", history=hist)
    b64 = UnevenBatchStepper(cases, precision="f64", history=hist)
    b32.step(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b32.step(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    b64.step(warmup + steps)
    torch.cuda.synchronize()
    err = 0.0
    for q in range(len(cases)):
        d32, d64 = b32.drag_history(q), b64.drag_history(q)
        err = max(err, float(np.max(np.abs(d32 - d64)) / np.max(np.abs(d64))))
    cells = sum(c["nz"] * c["nr"] for c in cases)
    del b64
    # the grouped layout, for comparison
    grouped = []
    for (nz, nr), specs in groups.items():
        args = [workloads.rigid_case_args(sp, workloads.case_inputs(sp, nz, nr)[1])
                for sp in specs]
        grouped.append(RigidBatchStepper(nz, nr, 1.0, args, r_len=0.5, precision="f32",
                                         history=hist))
    for g in grouped:
        g.step(warmup)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for g in grouped:
        g.step(steps)
    torch.cuda.synchronize()
    wall_g = time.perf_counter() - t1
    return dict(
        value=round(cells * steps / wall / 1e6, 2),
        unit="Mcell-steps/s",
        dtype="f32 fields/stencils/z-transforms; f64 radial solve, reductions, scalars",
        layout="one launch per pass over all cases (uneven grids)",
        cases=len(cases),
        groups={f"{nz}x{nr}": len(sp) for (nz, nr), sp in groups.items()},
        cells=cells,
        steps=steps,
        warmup=warmup,
        wall_s=round(wall, 4),
        drag_max_rel_err_vs_f64=err,
        grouped_value=round(cells * steps / wall_g / 1e6, 2),
        grouped_note="one batched stepper per grid, each on its own stream",
    )


def time_precision(grid_size_z=1024, no_cycles=20):
    """The precision form: an oscillating sphere in quiescent fluid (steady
    streaming; spheroid_streaming.py with aspect ratio 1) at nz x nr =
    1024 x 512, radius 0.1, amplitude 0.01, Womersley number 10, 20 periods.
    FP32 fields, stencils and z-transforms with the radial solve and force
    reduction in FP64, against the FP64 path over the same steps (the FP64
    path meets the oracle within 1e-10 in the test suite; the oracle itself
    would need ~an hour per run at this size)."""
    import torch

    from pyaxisymflow_amd.stepper import OscillatingSpheroidStepper

    def make(prec):
        return OscillatingSpheroidStepper(grid_size_z, womersley_square=100.0, radius=0.1,
                                          spheroid_AR=1.0, freq=16, no_cycles=no_cycles,
                                          precision=prec)

    # steps to the end time (the dt rule depends on the flow, so count them)
    probe = make("f64")
    while probe.time() < probe.tEnd * (1 - 1e-13):
        left = probe.tEnd - probe.time()
        probe.step(64 if left > 80 * probe.dt() else 1)
    n = probe.steps_done()
    del probe
    out = {}
    runs = {}
    for prec in ("f64", "f32"):
        st = make(prec)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st.stream)
        st.step(n)
        e1.record(st.stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        out[prec] = round(st.nr * st.nz * n / (ms / 1e3) / 1e6, 2)
        runs[prec] = st
    a, b = runs["f64"], runs["f32"]
    rel = lambda x, y: float(np.linalg.norm(x - y) / np.linalg.norm(y))
    ua, ub = a.velocity(), b.velocity()
    fa, fb = a.drag_history(), b.drag_history()
    return dict(
        value=out["f32"], unit="Mcell-steps/s", fp64_value=out["f64"],
        workload="oscillating sphere, steady streaming, 1024x512, 20 periods",
        steps=n, sim_time=b.time(),
        u_rel_l2_vs_f64=rel(np.hypot(*ub), np.hypot(*ua)),
        psi_rel_l2_vs_f64=rel(b.psi(), a.psi()),
        psi_avg_rel_l2_vs_f64=rel(b.avg_psi(), a.avg_psi()),
        drag_max_rel_err_vs_f64=float(np.abs(fb - fa).max() / np.abs(fa).max()),
        vs_oracle=precision_vs_oracle(),
    )


def precision_vs_oracle(grid_size_z=256, no_cycles=1):
    """The precision form against the FP64 oracle (the reference's loop
    restated on the CPU, exact Poisson route) over one whole period at
    256 x 128 (the oracle needs about an hour for 20 periods at 1024 x 512)."""
    from oracle.drivers import OscillatingSpheroidCase
    from pyaxisymflow_amd.stepper import OscillatingSpheroidStepper

    cpu = OscillatingSpheroidCase(grid_size_z, solver="exact", no_cycles=no_cycles)
    while cpu.t < cpu.tEnd:
        cpu.step()
    gpu = OscillatingSpheroidStepper(grid_size_z, precision="f32", no_cycles=no_cycles)
    gpu.step(cpu.it)
    rel = lambda x, y: float(np.linalg.norm(x - y) / np.linalg.norm(y))
    uz, ur = gpu.velocity()
    f, ref = gpu.drag_history(), np.array(cpu.force)
    return dict(workload=f"same driver at {grid_size_z}x{grid_size_z // 2}, {no_cycles} period, "
                         f"{cpu.it} steps",
                u_rel_l2=rel(np.hypot(uz, ur), np.hypot(cpu.u_z, cpu.u_r)),
                psi_rel_l2=rel(gpu.psi(), cpu.psi),
                drag_max_rel_err=float(np.abs(f - ref).max() / np.abs(ref).max()))


def time_implicit(sizes=((256, 128), (1024, 512), (2048, 1024)), reps=20):
    """Implicit Euler diffusion (kernels/implicit_diffusion_solver.py, the
    drop-in ImplicitEulerDiffusionStepper) on device-resident fields: one
    step = DST-I rows, radial tridiagonal solves, DST-I rows.  256 x 128 is
    the torus-streaming grid that uses it; dt = dx^2 / 4 / nu
    (flow_past_sphere.py:90).  The DST-I rows run as (nz + 1)-point DFTs
    (Stockham when nz + 1 factors into 2, 3 and 5, else chirp-z); the parity
    tests check the result against the oracle."""
    import torch

    from pyaxisymflow_amd.kernels.implicit_diffusion_solver import ImplicitEulerDiffusionStepper

    res = {}
    for nz, nr in sizes:
        dx, nu = 1.0 / nz, 1.5e-3
        dt = dx**2 / 4 / nu
        w0 = np.random.default_rng(0).standard_normal((nr, nz))
        st = ImplicitEulerDiffusionStepper(dt, nu, nr, nz, dx)
        w = torch.from_numpy(w0).cuda()
        st.step(w, dt)
        torch.cuda.synchronize()
        s = torch.cuda.current_stream()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            st.step(w, dt)
        e1.record(s)
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        res[f"{nz}x{nr}"] = dict(us_per_step=round(us, 2),
                                 mcell_steps_per_s=round(nr * nz / us, 2))
    return dict(unit="us per implicit step (drop-in, fields resident)", sizes=res,
                note="DST-I rows as (nz + 1)-point DFTs in LDS (chirp-z when nz + 1 is not "
                     "2,3,5-smooth)")


def time_potential(sizes=((256, 128), (2048, 1024)), reps=20):
    """Velocity potential (kernels/FastDiagonalisationPotentialSolver.py, the
    drop-in solver) plus compute_velocity_from_phi on device-resident fields:
    one solve = DCT-II rows, the mode-0 gauge solve, the radial solves,
    DCT-III rows; 256 x 128 is the torus-streaming grid that uses it."""
    import torch

    from pyaxisymflow_amd.kernels.compute_velocity_from_phi import compute_velocity_from_phi_unb
    from pyaxisymflow_amd.kernels.FastDiagonalisationPotentialSolver import (
        FastDiagonalisationPotentialSolver,
    )

    res = {}
    for nz, nr in sizes:
        dx = 1.0 / nz
        st = FastDiagonalisationPotentialSolver(nr, nz, dx)
        rhs = torch.from_numpy(np.random.default_rng(0).standard_normal((nr, nz))).cuda()
        phi, u, v = (torch.empty_like(rhs) for _ in range(3))
        st.solve(phi, rhs)
        compute_velocity_from_phi_unb(u, v, phi, dx)
        torch.cuda.synchronize()
        s = torch.cuda.current_stream()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            st.solve(phi, rhs)
            compute_velocity_from_phi_unb(u, v, phi, dx)
        e1.record(s)
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        res[f"{nz}x{nr}"] = dict(us_per_solve_and_velocity=round(us, 2),
                                 mcells_per_s=round(nr * nz / us, 2))
    return dict(unit="us per potential solve + velocity (drop-in, fields resident)", sizes=res)


def serving_precision_error(steps, nr=256, G=1.0):
    """One request of the serving set run to completion in the FP32-bulk
    and the FP64 soft batch: relative L2 of the FP32 fields against the
    FP64 ones (the FP64 batch matches the oracle to round-off)."""
    from pyaxisymflow_amd.stepper import SoftSphereBatch

    out = {}
    views = {}
    for prec in ("f32", "f64"):
        b = SoftSphereBatch(max_cases=1, precision=prec)
        c = b.add(grid_size_z=2 * nr, G=G)
        b.step(steps)
        e1, e2 = c.eta()
        views[prec] = dict(vorticity=c.vorticity().astype(np.float64),
                           psi=c.psi().astype(np.float64), eta1=e1, eta2=e2, phi=c.phi(),
                           t=c.time())
        del b
    for k in ("vorticity", "psi", "eta1", "eta2", "phi"):
        a, b = views["f32"][k], views["f64"][k]
        out[f"{k}_rel_l2"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out["time_rel_err"] = abs(views["f32"]["t"] - views["f64"]["t"]) / views["f64"]["t"]
    out["request"] = f"nr {nr}, nz {2 * nr}, G {G}, {steps} steps"
    return out


def time_serving(n_requests, steps, max_active=32, seed=0, precision="f32", open_loop=True):
    """The serving form: soft-sphere requests (pyaxisymflow_amd.serve:
    nr lognormal around 256 in [128, 1024]; G in [0.1, 10]), continuously
    batched (one launch per pass over the admitted requests).  A closed-loop
    pass (every request queued at t = 0) measures the capacity; the
    open-loop pass offers Poisson arrivals at 80% of it and reports the
    time-to-solution percentiles."""
    from pyaxisymflow_amd.serve import SoftServer, latency_summary, request_set

    srv = SoftServer(max_active=max_active, chunk=100, precision=precision)
    closed = request_set(n_requests, seed=seed, rate=1e9, steps=steps)
    t0 = time.perf_counter()
    recs = srv.run(closed)
    t_closed = time.perf_counter() - t0
    capacity = n_requests / t_closed
    # cell-steps actually done: a request evicted for a non-finite dt counts
    # only the steps it ran
    size = {r["id"]: r["nr"] * r["grid_size_z"] for r in closed}
    cells = sum(size[r["id"]] * r["steps"] for r in recs)
    failed = sum(1 for r in recs if r.get("failed"))
    if not open_loop:
        return dict(closed_loop_wall_s=round(t_closed, 3),
                    closed_loop_mcell_steps_per_s=round(cells / t_closed / 1e6, 2),
                    capacity_requests_per_s=round(capacity, 3),
                    lsq_tail_steps=int(sum(r["lsq_unconverged"] for r in recs)),
                    failed_requests=failed, precision=precision)
    rate = 0.8 * capacity
    open_set = request_set(n_requests, seed=seed, rate=rate, steps=steps)
    t0 = time.perf_counter()
    recs2 = srv.run(open_set)
    t_open = time.perf_counter() - t0
    s = latency_summary(recs2)
    return dict(
        p50_s=round(s["p50_s"], 4), p99_s=round(s["p99_s"], 4), mean_s=round(s["mean_s"], 4),
        unit="s (case time-to-solution: arrival to last step done)",
        requests=n_requests, steps_per_request=steps, max_active=max_active,
        nr_set=sorted({r["nr"] for r in open_set}),
        capacity_requests_per_s=round(capacity, 3), offered_rate_per_s=round(rate, 3),
        offered_load=0.8, open_loop_wall_s=round(t_open, 3),
        closed_loop_wall_s=round(t_closed, 3),
        closed_loop_mcell_steps_per_s=round(cells / t_closed / 1e6, 2),
        failed_requests=failed + s["failed"],
        # steps whose extrapolation loop outran the queued sweeps and was
        # finished by the one-workgroup tail (results unchanged)
        lsq_tail_steps=int(sum(r["lsq_unconverged"] for r in recs + recs2)),
        queued_lsq_sweeps=srv.last_batch_info[1] if srv.last_batch_info else None,
        precision=precision,
    )


def cpu_baseline(nz, nr, steps):
    """The oracle (numpy restatement of the reference path with the dense
    eig-route Poisson solve, BLAS on all cores) on the same workload."""
    from oracle.drivers import RigidBodyCase

    wl = workload(nz, nr)
    t0 = time.perf_counter()
    case = RigidBodyCase(nz, nr, wl["z_len"], wl["nu"], wl["brink_lam"], wl["char_func"],
                         U_0=wl["U_0"], r_sph=wl["r_sph"], vorticity=wl["vorticity"],
                         solver="eig")
    setup = time.perf_counter() - t0
    case.step()  # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        case.step()
    dt = time.perf_counter() - t0
    try:
        from threadpoolctl import threadpool_info

        blas = [(i.get("internal_api"), i.get("version"), i.get("num_threads"))
                for i in threadpool_info() if i.get("user_api") == "blas"]
    except Exception:
        blas = []
    cores = len(os.sched_getaffinity(0))
    threads = max([b[2] for b in blas if b[2]] or [cores])
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(
        value=nr * nz * steps / dt / 1e6,
        unit="Mcell-steps/s",
        cores=int(threads),
        kind="port",
        sample=(f"oracle (numpy, eig-route dense Poisson) at {nz}x{nr}: {steps} steps after 1 "
                f"warm-up, {dt:.2f} s; eig set-up {setup:.2f} s not included; "
                f"affinity {cores} cores ({model}); blas {blas}"),
    )


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nz", type=int, default=2048)
    ap.add_argument("--nr", type=int, default=1024)
    ap.add_argument("--stage-reps", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=200)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--large-steps", type=int, default=100,
                    help="steps of the 4096x2048 FP64 point (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--serve-steps", type=int, default=2000)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-implicit", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every case is independent and this build runs on "
              "one device; independent cases shard across GPUs without communication; "
              "run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    nz, nr = args.nz, args.nr
    ms, finite, cd, stages, copy_gbs, launch = time_gpu(nz, nr, args.steps, args.warmup,
                                                        args.stage_reps)
    if not finite:
        print("non-finite vorticity after the timed steps", file=sys.stderr)
        sys.exit(1)
    ms_step = ms / args.steps
    value = nr * nz * args.steps / (ms / 1e3) / 1e6
    dom = max(stages, key=stages.get)
    bpc = STAGE_BYTES_PER_CELL[dom]
    achieved = bpc * nr * nz / (stages[dom] * 1e-6) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"plain FP64 step, seeded 64-Gaussian vorticity + 8-sphere mask, {nz}x{nr}",
            "nz": nz, "nr": nr, "domain": "z in [0,2], r in [0,1]", "lambda": 1e4, "nu": 1.5e-3,
            "U": 1.0, "launch": launch, "last_Cd": cd,
            "stage_us": {k: round(v, 2) for k, v in stages.items()},
            "step_bytes_per_cell": sum(v for k, v in STAGE_BYTES_PER_CELL.items()
                                       if k not in ("body_mask", "implicit_diffusion")),
            "graded_bytes_per_cell": 88,
            "copy_kernel_GBs": round(copy_gbs, 1),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profile_traffic(dom, nz, nr),
            "traffic_unit": "B per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated)",
            "bytes_per_cell": bpc,
        },
    }
    forms = {}
    if args.large_steps > 0 and (nz, nr) != (4096, 2048):
        ms2, finite2, _, st2, _, _ = time_gpu(4096, 2048, args.large_steps, 8, 5)
        forms["fp64_4096x2048"] = dict(
            value=round(2048 * 4096 * args.large_steps / (ms2 / 1e3) / 1e6, 2),
            unit="Mcell-steps/s", ms_per_step=round(ms2 / args.large_steps, 5),
            steps=args.large_steps, finite=finite2,
            stage_us={k: round(v, 2) for k, v in st2.items()},
            note="same workload generator at nz x nr = 4096 x 2048 (z in [0,2], r in [0,1])")
    if not args.no_precision:
        forms["precision_fp32"] = time_precision()
    if not args.no_batched:
        forms["batched_fp32"] = time_batched(args.batched_steps, 2)
    if not args.no_implicit:
        forms["implicit_diffusion"] = time_implicit()
        forms["potential"] = time_potential()
    if not args.no_serve:
        sv = time_serving(args.serve_requests, args.serve_steps)
        sv["dtype"] = ("f32 flow fields/transforms/stencils; f64 radial elimination, reference "
                       "map, level set, reinit, extrapolation, scalars")
        sv["f64"] = time_serving(args.serve_requests, args.serve_steps, precision="f64",
                                 open_loop=False)
        sv["f32_vs_f64"] = serving_precision_error(args.serve_steps)
        forms["serving_soft"] = sv
    if forms:
        out["forms"] = forms
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(nz, nr, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
