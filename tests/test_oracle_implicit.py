"""The implicit Euler diffusion oracle (kernels/implicit_diffusion_solver.py)
pinned by what the mathematics fixes: the DST-I eigenpairs of the Dirichlet
P_z in closed form, the tridiagonal bands equal to the reference's dense
matrix bit for bit, both routes satisfying the discrete system by its
stencil (applied independently of either route's matrices), the nu dt = 0
identity, the dt check, and a planted operator mistake that the residual
catches."""
import numpy as np
import pytest
import scipy.fft as sfft

from oracle import implicit_diffusion as O

SIZES = [(64, 32), (128, 64), (90, 37)]  # (nz, nr)


def setup(nz, nr, seed=0):
    dx = 1.0 / nz
    nu = 1e-3
    dt = dx**2 / 4 / nu  # flow_past_sphere.py:90 (diffusion_dt_limit)
    rng = np.random.default_rng(seed)
    return dx, nu, dt, rng.standard_normal((nr, nz))


def test_dst1_diagonalises_dirichlet_Pz():
    nz = 48
    dx = 1.0 / nz
    P_z, _, _ = O.fdm_matrices(8, nz, dx)
    S = sfft.dst(np.eye(nz), type=1, axis=0, norm="ortho")  # columns = basis vectors
    lam = O.dirichlet_z_eigenvalues(nz, dx)
    assert np.max(np.abs(P_z @ S - S * lam[None, :])) * dx**2 < 1e-12
    assert np.allclose(S @ S, np.eye(nz), atol=1e-13)  # symmetric and orthogonal
    assert np.allclose(np.sort(np.linalg.eigvalsh(P_z)), np.sort(lam), rtol=0, atol=1e-11 / dx**2)


def test_radial_bands_bitwise_match_dense():
    for nr in (5, 37, 128):
        dx = 1.0 / 256
        _, P_r, N_r = O.fdm_matrices(nr, 16, dx)
        A = P_r + N_r
        sub, diag, sup = O.radial_bands(nr, dx)
        assert np.array_equal(np.diag(A), diag)
        assert np.array_equal(np.diag(A, -1), sub[1:])
        assert np.array_equal(np.diag(A, 1), sup[:-1])
        assert np.count_nonzero(A - np.diag(np.diag(A)) - np.diag(np.diag(A, 1), 1)
                                - np.diag(np.diag(A, -1), -1)) == 0


@pytest.mark.parametrize("nz,nr", SIZES)
def test_both_routes_solve_the_discrete_system(nz, nr):
    dx, nu, dt, w = setup(nz, nr)
    exact = O.exact_implicit_diffusion(w, nu * dt, dx)
    assert O.implicit_diffusion_residual(exact, w, nu * dt, dx) < 1e-13
    eig = w.copy()
    O.ImplicitEulerDiffusionStepper(dt, nu, nr, nz, dx).step(eig, dt)
    assert O.implicit_diffusion_residual(eig, w, nu * dt, dx) < 1e-11
    assert np.linalg.norm(eig - exact) / np.linalg.norm(exact) < 1e-11


def test_zero_viscosity_is_identity_and_dt_is_checked():
    nz, nr = 32, 16
    dx, nu, dt, w = setup(nz, nr)
    assert np.allclose(O.exact_implicit_diffusion(w, 0.0, dx), w, rtol=0, atol=1e-14)
    st = O.ImplicitEulerDiffusionStepper(dt, 0.0, nr, nz, dx)
    a = w.copy()
    st.step(a, dt)
    assert np.allclose(a, w, rtol=0, atol=1e-12)
    with pytest.raises(ValueError):
        st.step(a, 2 * dt)


def test_diffusion_contracts_and_planted_sign_error_is_caught():
    nz, nr = 64, 32
    dx, nu, dt, w = setup(nz, nr, seed=3)
    out = O.exact_implicit_diffusion(w, 10 * nu * dt, dx)
    # (I - nu dt L)^{-1} with L's spectrum <= 0 shrinks the field
    assert np.max(np.abs(out)) < np.max(np.abs(w))
    # planted mistake: the 1/r first-derivative term with the wrong sign
    sub, diag, sup = O.radial_bands(nr, dx)
    orig = O.radial_bands
    try:
        O.radial_bands = lambda n, d: (sup, diag, sub)
        wrong = O.exact_implicit_diffusion(w, nu * dt, dx)
    finally:
        O.radial_bands = orig
    assert O.implicit_diffusion_residual(wrong, w, nu * dt, dx) > 1e-3
