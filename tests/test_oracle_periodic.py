"""The periodic-z Stokes oracle routes (FastDiagonalisationStokesSolver.py:88-97:
Neumann or Dirichlet at r_max, periodic P_z) pinned by the DFT eigenpairs of
the circulant P_z, the exact route's stencil residual, agreement with the
reference's eig route where that route is itself accurate, and the packed
real-DFT layout the GPU uses (its per-entry eigenvalues and its inverse)."""
import numpy as np
import pytest
import scipy.fft as sfft

from oracle.stokes import (
    DIRICHLET_R_PERIODIC_Z, NEUMANN_R_PERIODIC_Z, FastDiagonalisationStokesSolver,
    exact_stokes_solve, fdm_matrices, stokes_residual, z_eigenvalues,
)

BCS = [NEUMANN_R_PERIODIC_Z, DIRICHLET_R_PERIODIC_Z]


@pytest.mark.parametrize("bc", BCS)
def test_dft_diagonalises_periodic_Pz(bc):
    nz = 24
    dx = 1.0 / nz
    _, P_z, _, _ = fdm_matrices(6, nz, dx, bc)
    assert P_z[0, -1] == P_z[0, 1] and P_z[-1, 0] == P_z[-1, -2]
    lam = z_eigenvalues(nz, dx, bc)
    F = np.exp(2j * np.pi * np.outer(np.arange(nz), np.arange(nz)) / nz)  # columns: modes
    assert np.max(np.abs(P_z @ F - F * lam[None, :])) * dx**2 < 1e-12


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("nz,nr", [(64, 32), (100, 37)])
def test_exact_periodic_route_residual(bc, nz, nr):
    dx = 1.0 / nz
    w = np.random.default_rng(0).standard_normal((nr, nz))
    assert stokes_residual(exact_stokes_solve(w, dx, bc), w, dx, bc) < 1e-13


@pytest.mark.parametrize("bc", BCS)
def test_eig_route_agrees_where_accurate(bc):
    """At 100 x 37 the reference's eig route solves its system to 1e-13 and
    matches the exact route (at power-of-two grids its degenerate periodic
    eigenspaces can leave it 3e-3 off in the Dirichlet case, by its own
    residual, so parity is pinned on the exact route there)."""
    nz, nr = 100, 37
    dx = 1.0 / nz
    w = np.random.default_rng(1).standard_normal((nr, nz))
    p = np.zeros_like(w)
    FastDiagonalisationStokesSolver(nr, nz, dx, bc_type=bc).solve(p, w)
    e = exact_stokes_solve(w, dx, bc)
    assert stokes_residual(p, w, dx, bc) < 1e-11
    assert np.linalg.norm(p - e) / np.linalg.norm(e) < 1e-11


def test_packed_real_dft_layout():
    """[X_0, Re X_1, Im X_1, ..., X_M] with M = nz / 2 holds the real DFT; the
    per-entry eigenvalues are those of the entry's mode."""
    nz = 12
    M = nz // 2
    x = np.random.default_rng(2).standard_normal(nz)
    X = np.fft.rfft(x)
    packed = np.concatenate([[X[0].real], np.column_stack([X[1:M].real, X[1:M].imag]).ravel(),
                             [X[M].real]])
    assert packed.shape == (nz,)
    assert abs(X[0].imag) < 1e-14 and abs(X[M].imag) < 1e-14
    modes = np.array([0] + [k for k in range(1, M) for _ in (0, 1)] + [M])
    lam = z_eigenvalues(nz, 1.0, NEUMANN_R_PERIODIC_Z)
    assert np.allclose(lam[modes], lam[(nz - modes) % nz])  # k and N - k share an eigenvalue
    back = np.zeros(M + 1, complex)
    back[0], back[M] = packed[0], packed[-1]
    back[1:M] = packed[1:-1:2] + 1j * packed[2:-1:2]
    assert np.allclose(np.fft.irfft(back, nz), x, atol=1e-14)


def test_ghost_exchange_restores_periodic_fields():
    """A field sampled from a z-periodic function of period (nz - 2g) dx is
    its own exchange; the shifted exchange restores the identity reference
    map eta = z (its ghosts are one period away)."""
    from oracle.periodic import gen_periodic_boundary_ghost_comm, gen_periodic_boundary_ghost_comm_eta

    nz, nr, g = 40, 6, 2
    dx = 0.05
    z = (np.arange(nz) + 0.5) * dx - g * dx
    period = (nz - 2 * g) * dx
    f = np.sin(2 * np.pi * z / period)[None, :] * np.arange(1, nr + 1)[:, None]
    h = f.copy()
    h[:, :g] = 7.0
    h[:, -g:] = -3.0
    gen_periodic_boundary_ghost_comm(g)(h)
    assert np.allclose(h, f, rtol=0, atol=1e-13)
    eta = np.tile(z, (nr, 1))
    e2 = eta.copy()
    e2[:, :g] = 0.0
    e2[:, -g:] = 0.0
    Z_max = nz * dx
    gen_periodic_boundary_ghost_comm_eta(g, Z_max, dx)(e2)
    assert np.allclose(e2, eta, rtol=0, atol=1e-13)


def test_periodic_rk2_preserves_z_periodicity_and_planted_error_is_caught():
    """diffusion_RK2_periodic on a z-periodic field keeps its interior
    periodic; skipping the exchange between the stages (a planted mistake)
    breaks that at the seam."""
    from oracle import kernels as K
    from oracle.periodic import diffusion_RK2_periodic, gen_periodic_boundary_ghost_comm

    nz, nr, g = 36, 10, 1
    dx = 1.0 / 32
    z = (np.arange(nz) + 0.5 - g) * dx
    r = (np.arange(nr) + 0.5) * dx
    Z, R = np.meshgrid(z, r)
    period = (nz - 2 * g) * dx
    w = np.cos(2 * np.pi * Z / period) * R * np.exp(-R)
    comm = gen_periodic_boundary_ghost_comm(g)
    a, ta = w.copy(), np.zeros_like(w)
    diffusion_RK2_periodic(a, ta, R, 0.01, 1e-3, dx, comm)
    # interior columns one period apart agree
    assert np.allclose(a[1:-1, g + 1], a[1:-1, nz - g + 1 - (nz - 2 * g)], atol=1e-14)
    b, tb = w.copy(), np.zeros_like(w)
    comm(b)
    K._rk2_stage1(b, tb, R, 0.01, 1e-3, dx)
    tb[:, :g] = 0.0  # the planted mistake: stale ghosts in the half step
    K._rk2_stage2(b, tb, R, 0.01, 1e-3, dx)
    assert np.max(np.abs(b - a)) > 1e-6
