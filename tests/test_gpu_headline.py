"""GPU parity of the rigid-body step at the kernel-scale sizes the throughput
figure is measured on (row A16 at 2048 x 1024 and 4096 x 2048).

Workload: the seeded kernel-scale field (pyaxisymflow_amd.workloads:
64 Gaussians, 8 on-axis spheres, lambda = 1e4, nu = 1.5e-3, U = 1, no
ramp), stepped by RigidBodyStepper and by the oracle's restatement of
flow_past_sphere.py:107-188 (oracle.drivers.RigidBodyCase).

Bar (FP64): vorticity, psi and the drag history within 1e-10 relative of
the oracle's exact Poisson route (DCT-II + banded radial solve).  The
reference's own eig route (FastDiagonalisationStokesSolver.py:99-156) is
itself inexact at this size (about 1.3e-10 relative to the exact discrete
solution), so against it the bar is 1e-9 and the error is printed.

At nr >= 1024 the radial pass's layout (segment length L, modes per
workgroup K, coefficients in LDS or global memory) is picked by a timed
search; every layout the search can pick is pinned in turn through
AXS_TRIDIAG_LK and held to the same bar."""
import numpy as np
import pytest

from oracle.drivers import RigidBodyCase
from pyaxisymflow_amd.workloads import kernel_scale_workload

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


_CACHE = {}


def oracle_run(nz, nr, steps, solver):
    key = (nz, nr, steps, solver)
    if key not in _CACHE:
        wl = kernel_scale_workload(nz, nr)
        cpu = RigidBodyCase(nz, nr, wl["z_len"], wl["nu"], wl["brink_lam"], wl["char_func"],
                            U_0=wl["U_0"], T_ramp=None, r_sph=wl["r_sph"],
                            vorticity=wl["vorticity"], solver=solver)
        for _ in range(steps):
            cpu.step()
        _CACHE[key] = (wl, cpu)
    return _CACHE[key]


def gpu_run(wl, nz, nr, steps):
    from pyaxisymflow_amd.stepper import RigidBodyStepper

    gpu = RigidBodyStepper(nz, nr, wl["z_len"], wl["nu"], wl["brink_lam"], wl["char_func"],
                           U_0=wl["U_0"], T_ramp=None, r_sph=wl["r_sph"],
                           vorticity=wl["vorticity"], history=steps + 8)
    gpu.step(steps)
    return gpu


def check(gpu, cpu, bar, label):
    ew = rel_l2(gpu.vorticity(), cpu.vorticity)
    ep = rel_l2(gpu.psi(), cpu.psi)
    d = gpu.drag_history()
    ref = np.array(cpu.drag)
    ed = float(np.max(np.abs(d - ref) / np.maximum(np.abs(ref), 1e-300)))
    print(f"{label}: omega {ew:.2e} psi {ep:.2e} Cd {ed:.2e}", flush=True)
    assert len(d) == len(ref)
    assert ew < bar and ep < bar and ed < bar, (ew, ep, ed)


# (L, K, G): the headline's usual pick first, then other admissible layouts
# of the search at nr = 1024 (P = nr / L segments, 64 <= K P <= 1024)
LAYOUTS = [None, "16,8,0", "16,16,1", "8,4,0", "8,8,1", "4,2,0", "4,4,1"]


@pytest.mark.parametrize("pin", LAYOUTS)
def test_2048x1024_three_steps_vs_exact(pin, monkeypatch):
    if pin is None:
        monkeypatch.delenv("AXS_TRIDIAG_LK", raising=False)
    else:
        monkeypatch.setenv("AXS_TRIDIAG_LK", pin)
    wl, cpu = oracle_run(2048, 1024, 3, "exact")
    gpu = gpu_run(wl, 2048, 1024, 3)
    check(gpu, cpu, 1e-10, f"2048x1024 layout {pin or 'searched'} vs exact")


def test_2048x1024_three_steps_vs_eig_route():
    wl, cpu = oracle_run(2048, 1024, 3, "eig")
    gpu = gpu_run(wl, 2048, 1024, 3)
    check(gpu, cpu, 1e-9, "2048x1024 vs the eig route")


def test_4096x2048_two_steps_vs_exact():
    wl, cpu = oracle_run(4096, 2048, 2, "exact")
    gpu = gpu_run(wl, 4096, 2048, 2)
    check(gpu, cpu, 1e-10, "4096x2048 vs exact")
