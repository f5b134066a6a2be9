"""GPU parity of the fused rigid-body step (row A16) against the oracle's
restatement of the driver loop (flow_past_sphere.py:107-188).

Workload: flow past a rigid sphere, nz x nr = 256 x 128, z in [0, 2],
r in [0, 1], sphere radius 0.075 at z = 0.5, U = 1 with the reference's sine
ramp (T_ramp = 20 r / U), Re = 100 on the diameter, lambda = 1e4, dt from
min(CFL 0.1, 0.9 dx^2 / (4 nu)).  FP64 bar: fields within 1e-10 relative L2
and drag within 1e-10 relative of the oracle (against both the exact
Poisson route and the reference's eig route)."""
import numpy as np
import pytest

from oracle.drivers import RigidBodyCase, gaussian_vorticity, grid
from oracle.kernels import smooth_Heaviside

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def sphere_case(nz=256, nr=128, z_len=2.0, seed=None):
    dx, z, r, Z, R = grid(nz, nr, z_len)
    r_sph, U = 0.075, 1.0
    nu = U * 2 * r_sph / 100
    phi = r_sph - np.sqrt((Z - 0.5) ** 2 + R**2)
    chi = np.zeros_like(phi)
    smooth_Heaviside(chi, phi, dx * 2**0.5)
    w0 = None
    if seed is not None:
        w0 = gaussian_vorticity(Z, R, 16, seed, 0.02, 0.1, z_len, nr * dx)
    return dict(nz=nz, nr=nr, z_len=z_len, nu=nu, brink_lam=1e4, char_func=chi, U_0=U,
                T_ramp=20 * r_sph / U, r_sph=r_sph, vorticity=w0)


def run_pair(case, steps, solver="exact", use_graph=True):
    from pyaxisymflow_amd.stepper import RigidBodyStepper

    cpu = RigidBodyCase(case["nz"], case["nr"], case["z_len"], case["nu"], case["brink_lam"],
                        case["char_func"], U_0=case["U_0"], T_ramp=case["T_ramp"],
                        r_sph=case["r_sph"], vorticity=case["vorticity"], solver=solver)
    gpu = RigidBodyStepper(case["nz"], case["nr"], case["z_len"], case["nu"], case["brink_lam"],
                           case["char_func"], U_0=case["U_0"], T_ramp=case["T_ramp"],
                           r_sph=case["r_sph"], vorticity=case["vorticity"], use_graph=use_graph)
    for _ in range(steps):
        cpu.step()
    gpu.step(steps)
    return cpu, gpu


def test_one_step_seeded_field_every_stage():
    """One full step on the seeded 16-Gaussian field (the one-step kernel
    check): vorticity, psi, penalised velocity, dt and Cd."""
    case = sphere_case(seed=0)
    cpu, gpu = run_pair(case, 1)
    # dt = CFL dx / max|u| inherits the Poisson solve's round-off (1e-13)
    assert abs(gpu.dt() - cpu.dts[0]) <= 1e-12 * cpu.dts[0]
    assert rel_l2(gpu.psi(), cpu.psi) < 1e-12
    uz, ur = gpu.velocity()
    assert rel_l2(uz, cpu.u_z) < 1e-12 and rel_l2(ur, cpu.u_r) < 1e-12
    assert rel_l2(gpu.vorticity(), cpu.vorticity) < 1e-12
    d = gpu.drag_history()
    assert len(d) == 1 and abs(d[0] - cpu.drag[0]) <= 1e-12 * max(1.0, abs(cpu.drag[0]))
    assert abs(gpu.time() - cpu.t) <= 1e-12 * cpu.t


@pytest.mark.parametrize("solver", ["exact", "eig"])
def test_flow_past_sphere_500_steps(solver):
    """The plain-method configuration in full: 500 steps."""
    case = sphere_case()
    cpu, gpu = run_pair(case, 500, solver=solver)
    assert gpu.steps_done() == 500
    assert rel_l2(gpu.vorticity(), cpu.vorticity) < 1e-10
    assert rel_l2(gpu.psi(), cpu.psi) < 1e-10
    d = gpu.drag_history()
    np.testing.assert_allclose(d, np.array(cpu.drag), rtol=1e-10, atol=1e-12)
    assert abs(gpu.time() - cpu.t) <= 1e-12 * cpu.t


def test_flow_past_sphere_dx_not_power_of_two():
    """nz x nr = 240 x 120 on z in [0, 2]: dx = 1/120, so the fused step takes
    its general quotient path (the power-of-two grids above take products)."""
    case = sphere_case(nz=240, nr=120, seed=1)
    cpu, gpu = run_pair(case, 40)
    assert rel_l2(gpu.vorticity(), cpu.vorticity) < 1e-10
    assert rel_l2(gpu.psi(), cpu.psi) < 1e-10
    np.testing.assert_allclose(gpu.drag_history(), np.array(cpu.drag), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("nz,nr", [(256, 40), (384, 40), (512, 40), (768, 40), (1024, 40),
                                   (2048, 40), (4096, 40)])
def test_register_fft_row_passes_vs_oracle(nz, nr):
    """The tapered forward row pass and the inverse row pass of the register
    FFTs (nz = 256 to 2048: one wavefront per row; nz = 4096: two, joined
    by a radix-2 step) inside the step: 5 steps against the oracle."""
    case = sphere_case(nz=nz, nr=nr, seed=3)
    cpu, gpu = run_pair(case, 5)
    assert rel_l2(gpu.vorticity(), cpu.vorticity) < 1e-10
    assert rel_l2(gpu.psi(), cpu.psi) < 1e-10
    np.testing.assert_allclose(gpu.drag_history(), np.array(cpu.drag), rtol=1e-10, atol=1e-12)


def test_pinned_radial_solve_configuration(monkeypatch):
    """AXS_TRIDIAG_LK pins the tridiagonal pass's segment length and modes per
    workgroup: runs with the same pin are bit-identical, every pin meets the
    FP64 bar against the oracle, and a pin the kernel cannot run is refused."""
    from pyaxisymflow_amd.stepper import RigidBodyStepper

    case = sphere_case(seed=2)
    keys = ("nz", "nr", "z_len", "nu", "brink_lam", "char_func")
    kw = dict(U_0=case["U_0"], T_ramp=case["T_ramp"], r_sph=case["r_sph"],
              vorticity=case["vorticity"])
    cpu = RigidBodyCase(*(case[k] for k in keys), solver="exact", **kw)
    for _ in range(20):
        cpu.step()
    seen = {}
    for pin in ("8,2", "16,4", "4,1", "8,2"):
        monkeypatch.setenv("AXS_TRIDIAG_LK", pin)
        gpu = RigidBodyStepper(*(case[k] for k in keys), **kw)
        gpu.step(20)
        w = gpu.vorticity()
        assert rel_l2(w, cpu.vorticity) < 1e-10, pin
        if pin in seen:
            assert np.array_equal(seen[pin], w)
        seen[pin] = w
    monkeypatch.setenv("AXS_TRIDIAG_LK", "8,4096")  # 4096 x 16 threads per workgroup
    with pytest.raises(RuntimeError):
        RigidBodyStepper(*(case[k] for k in keys), **kw)


def test_graph_replay_matches_eager():
    """A captured step replayed 20 times equals 20 eagerly launched steps bit
    for bit (no memset nodes; every replay re-reads the device scalars)."""
    case = sphere_case(seed=1)
    from pyaxisymflow_amd.stepper import RigidBodyStepper

    kw = dict(U_0=case["U_0"], T_ramp=case["T_ramp"], r_sph=case["r_sph"],
              vorticity=case["vorticity"])
    args = (case["nz"], case["nr"], case["z_len"], case["nu"], case["brink_lam"], case["char_func"])
    a = RigidBodyStepper(*args, use_graph=True, **kw)
    b = RigidBodyStepper(*args, use_graph=False, **kw)
    for _ in range(4):
        a.step(5)
        a.synchronize()
        b.step(5)
    assert np.array_equal(a.vorticity(), b.vorticity())
    # multi-step graphs (16 steps per launch) plus single-step remainders
    a.step(37)
    b.step(37)
    assert a.steps_done() == b.steps_done() == 57
    assert np.array_equal(a.vorticity(), b.vorticity())
    assert np.array_equal(a.drag_history(), b.drag_history())
    assert np.array_equal(a.scalars(), b.scalars())


def test_chunk_graphs_after_an_odd_step_count():
    """A run of n steps launches n // 16 graphs of 16 steps and one graph per
    set bit of n % 16 (the last one completing the run): step(5) is 4 + 1
    and leaves the vorticity pair at the odd parity, step(20) is 16 + 4 from
    that parity, step(32) is 16 + 16; every run stays bit-equal to eager
    stepping."""
    case = sphere_case(seed=4)
    from pyaxisymflow_amd.stepper import RigidBodyStepper

    kw = dict(U_0=case["U_0"], T_ramp=case["T_ramp"], r_sph=case["r_sph"],
              vorticity=case["vorticity"])
    args = (case["nz"], case["nr"], case["z_len"], case["nu"], case["brink_lam"], case["char_func"])
    a = RigidBodyStepper(*args, use_graph=True, **kw)
    b = RigidBodyStepper(*args, use_graph=False, **kw)
    a.step(5)
    b.step(5)
    assert a.graph_launches() == (0, 2)
    a.step(20)
    b.step(20)
    assert a.graph_launches() == (1, 3)
    a.step(32)
    b.step(32)
    assert a.graph_launches() == (3, 3)
    assert a.steps_done() == b.steps_done() == 57
    assert np.array_equal(a.vorticity(), b.vorticity())
    assert np.array_equal(a.psi(), b.psi())
    assert np.array_equal(a.drag_history(), b.drag_history())
    assert np.array_equal(a.scalars(), b.scalars())


def test_oscillating_body_penalisation_velocity():
    """spheroid_streaming.py:216-226: the body's penalisation velocity
    U_0 cos(omega t) replaces the free stream (no ramp, no added U)."""
    from pyaxisymflow_amd.stepper import RigidBodyStepper

    nz, nr, z_len = 128, 64, 1.0
    dx, z, r, Z, R = grid(nz, nr, z_len)
    phi = 0.1 - np.sqrt((Z - 0.5) ** 2 + R**2)
    chi = np.zeros_like(phi)
    smooth_Heaviside(chi, phi, dx * 2**0.5)
    U_amp, omega, nu = 0.1 * 0.1 * 2 * np.pi * 16, 2 * np.pi * 16, 1e-3
    cpu = RigidBodyCase(nz, nr, z_len, nu, 1e4, chi, oscillation=(U_amp, omega), solver="exact")
    gpu = RigidBodyStepper(nz, nr, z_len, nu, 1e4, chi, oscillation=(U_amp, omega))
    for _ in range(30):
        cpu.step()
    gpu.step(30)
    assert rel_l2(gpu.vorticity(), cpu.vorticity) < 1e-10
    assert abs(gpu.time() - cpu.t) <= 1e-12 * cpu.t
