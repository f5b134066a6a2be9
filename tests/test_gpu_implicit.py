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
