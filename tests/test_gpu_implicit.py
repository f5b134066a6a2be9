"""GPU parity of the implicit Euler diffusion drop-in
(kernels/implicit_diffusion_solver.py:6-143) through the C ABI.

Bars (FP64): within 1e-12 relative L2 of the exact discrete solution and
1e-10 of the reference's eig route (itself inexact by the conditioning of
its non-normal radial eigenvectors), and the discrete system satisfied by
its stencil to 1e-12 at every size, including 2048 x 1024 where the CPU
routes are not run."""
import numpy as np
import pytest

from oracle import implicit_diffusion as O

pytestmark = pytest.mark.gpu

# (nz, nr): the torus-streaming grid, a tile-ragged grid, a tall one
SIZES = [(256, 128), (100, 37), (128, 200)]


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def case(nz, nr, seed=0):
    dx = 1.0 / nz
    nu = 1.5e-3
    dt = dx**2 / 4 / nu  # flow_past_sphere.py:90 (diffusion_dt_limit)
    rng = np.random.default_rng(seed)
    return dx, nu, dt, rng.standard_normal((nr, nz))


@pytest.mark.parametrize("nz,nr", SIZES)
def test_implicit_diffusion_matches_oracle(nz, nr):
    from pyaxisymflow_amd.kernels.implicit_diffusion_solver import ImplicitEulerDiffusionStepper

    dx, nu, dt, w = case(nz, nr)
    exact = O.exact_implicit_diffusion(w, nu * dt, dx)
    eig = w.copy()
    O.ImplicitEulerDiffusionStepper(dt, nu, nr, nz, dx).step(eig, dt)
    gpu = w.copy()
    st = ImplicitEulerDiffusionStepper(time_step=dt, kinematic_viscosity=nu, grid_size_r=nr,
                                       grid_size_z=nz, dx=dx)
    assert st.step(vorticity_field=gpu, dt=dt) is None
    e_exact, e_eig = rel_l2(gpu, exact), rel_l2(gpu, eig)
    print(f"implicit {nz}x{nr}: vs exact {e_exact:.2e}, vs eig route {e_eig:.2e}")
    assert e_exact < 1e-12
    assert e_eig < 1e-10
    assert O.implicit_diffusion_residual(gpu, w, nu * dt, dx) < 1e-12


def test_implicit_diffusion_repeated_steps_and_dt_check():
    from pyaxisymflow_amd.kernels.implicit_diffusion_solver import ImplicitEulerDiffusionStepper

    nz, nr = 256, 128
    dx, nu, dt, w = case(nz, nr, seed=1)
    st = ImplicitEulerDiffusionStepper(dt, nu, nr, nz, dx)
    gpu, ref = w.copy(), w.copy()
    for _ in range(10):
        st.step(gpu, dt)
        ref = O.exact_implicit_diffusion(ref, nu * dt, dx)
    assert rel_l2(gpu, ref) < 1e-12
    with pytest.raises(ValueError):
        st.step(gpu, dt * 1.5)


def test_implicit_diffusion_torch_in_place_large():
    import torch

    from pyaxisymflow_amd.kernels.implicit_diffusion_solver import ImplicitEulerDiffusionStepper

    nz, nr = 2048, 1024
    dx, nu, dt, w = case(nz, nr, seed=2)
    st = ImplicitEulerDiffusionStepper(dt, nu, nr, nz, dx)
    t = torch.from_numpy(w).cuda()
    ptr = t.data_ptr()
    st.step(t, dt)
    torch.cuda.synchronize()
    assert t.data_ptr() == ptr
    out = t.cpu().numpy()
    res = O.implicit_diffusion_residual(out, w, nu * dt, dx)
    print(f"implicit 2048x1024 residual {res:.2e}")
    assert res < 1e-12


def implicit_sphere_case(nz=256, nr=128, z_len=2.0, nu=0.02):
    """Flow past a sphere with the implicit-diffusion branch of
    flow_past_sphere.py (:90-98, 146-148, 179-181): fixed dt = dx^2 / 4 / nu;
    nu = 0.02 (Re = 7.5) keeps that dt under the advective limit."""
    from oracle.drivers import gaussian_vorticity, grid
    from oracle.kernels import smooth_Heaviside

    dx, z, r, Z, R = grid(nz, nr, z_len)
    phi = 0.075 - np.sqrt((Z - 0.5) ** 2 + R**2)
    chi = np.zeros_like(phi)
    smooth_Heaviside(chi, phi, dx * 2**0.5)
    w0 = gaussian_vorticity(Z, R, 16, 0, 0.02, 0.1, z_len, nr * dx)
    return (nz, nr, z_len, nu, 1e4, chi), dict(U_0=1.0, T_ramp=1.5, r_sph=0.075, vorticity=w0)


@pytest.mark.parametrize("solver", ["exact", "eig"])
def test_stepper_implicit_diffusion_matches_oracle_driver(solver):
    from oracle.drivers import RigidBodyCase
    from pyaxisymflow_amd.stepper import RigidBodyStepper

    args, kw = implicit_sphere_case()
    steps = 30
    cpu = RigidBodyCase(*args, solver=solver, implicit_diffusion=True, **kw)
    for _ in range(steps):
        cpu.step()
    gpu = RigidBodyStepper(*args, implicit_diffusion=True, **kw)
    gpu.step(steps)
    e_w = rel_l2(gpu.vorticity(), cpu.vorticity)
    cd = gpu.drag_history()
    e_cd = np.max(np.abs(cd[:steps] - np.array(cpu.drag))) / np.max(np.abs(cpu.drag))
    print(f"implicit stepper ({solver}): omega {e_w:.2e}, Cd {e_cd:.2e}, dt {cpu.dts[0]:.3e}")
    assert abs(gpu.time() - cpu.t) <= 1e-13 * cpu.t
    assert e_w < 1e-10
    assert e_cd < 1e-10


def test_stepper_implicit_graph_replay_matches_eager():
    from pyaxisymflow_amd.stepper import RigidBodyStepper

    args, kw = implicit_sphere_case(nz=128, nr=64)
    a = RigidBodyStepper(*args, implicit_diffusion=True, use_graph=True, **kw)
    b = RigidBodyStepper(*args, implicit_diffusion=True, use_graph=False, **kw)
    for n in (16, 3, 16):
        a.step(n)
        b.step(n)
        assert np.array_equal(a.vorticity(), b.vorticity())
    assert np.array_equal(a.drag_history()[:35], b.drag_history()[:35])
