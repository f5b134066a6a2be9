"""numpy restatement of the reference kernels on the hot path (test oracle only).

Every function follows the reference's array expressions in the same order,
so the FP64 results agree with the reference to round-off.  Citations are
``path:line`` inside the reference tree (pyaxisymflow v0.0.1).

Fields are C-order float64 arrays of shape (nr, nz): axis 0 = r (rows),
axis 1 = z (columns).
"""
import numpy as np

# ---------------------------------------------------------------------------
# A1  kill_boundary_vorticity_sine.py:4-27
# ---------------------------------------------------------------------------


def kill_boundary_vorticity_sine_z(vorticity, Z, width, dx):
    """kill_boundary_vorticity_sine.py:4-14.  The literal ``1`` is the
    reference's assumed domain end in z (kept as is)."""
    vorticity[:, :width] = np.sin(
        np.pi * (Z[:, :width] - 0.5 * dx) / 2 / (width - 1) / dx
    ) * vorticity[:, (width - 1)].reshape(-1, 1)
    vorticity[:, -width:] = np.sin(
        np.pi * (1 - Z[:, -width:] - 0.5 * dx) / 2 / (width - 1) / dx
    ) * vorticity[:, -width].reshape(-1, 1)


def kill_boundary_vorticity_sine_r(vorticity, R, width, dx):
    """kill_boundary_vorticity_sine.py:17-27 (literal ``1`` as r end)."""
    vorticity[-width:, :] = (
        np.sin(np.pi * (1 - R[-width:, :] - 0.5 * dx) / 2 / (width - 1) / dx)
        * vorticity[-width, :]
    )
    vorticity[0, :] = 0.0


# ---------------------------------------------------------------------------
# A3  compute_velocity_from_psi.py:4-17
# ---------------------------------------------------------------------------


def compute_velocity_from_psi_unb(u_z, u_r, psi, R, dx):
    """compute_velocity_from_psi.py:4-17"""
    u_z[1:-1, :] = (psi[2:, :] - psi[:-2, :]) / (2 * dx) / R[1:-1, :]
    u_z[0, :] = (-psi[2, :] + 4 * psi[1, :] - 3 * psi[0, :]) / (2 * dx) / R[0, :]
    u_z[-1, :] = (psi[-3, :] - 4 * psi[-2, :] + 3 * psi[-1, :]) / (2 * dx) / R[-1, :]
    u_r[:, 1:-1] = -(psi[:, 2:] - psi[:, :-2]) / (2 * dx) / R[:, 1:-1]
    u_r[:, 0] = -(-psi[:, 2] + 4 * psi[:, 1] - 3 * psi[:, 0]) / (2 * dx) / R[:, 0]
    u_r[:, -1] = -(psi[:, -3] - 4 * psi[:, -2] + 3 * psi[:, -1]) / (2 * dx) / R[:, -1]


# ---------------------------------------------------------------------------
# A4  compute_vorticity_from_velocity.py:4-13
# ---------------------------------------------------------------------------


def compute_vorticity_from_velocity_unb(vort, u_z, u_r, dx):
    """compute_vorticity_from_velocity.py:4-13 (edges untouched)"""
    vort[1:-1, 1:-1] = (u_r[1:-1, 2:] - u_r[1:-1, :-2]) / (2 * dx) - (
        u_z[2:, 1:-1] - u_z[:-2, 1:-1]
    ) / (2 * dx)


# ---------------------------------------------------------------------------
# A5  brinkmann_penalize.py:4-16
# ---------------------------------------------------------------------------


def brinkmann_penalize(
    lam, dt, char_func, U_z, U_r, grid_u_z, grid_u_r, penalized_u_z, penalized_u_r
):
    """brinkmann_penalize.py:4-16"""
    penalized_u_z[...] = (grid_u_z + lam * dt * char_func * U_z) / (
        1 + lam * dt * char_func
    )
    penalized_u_r[...] = (grid_u_r + lam * dt * char_func * U_r) / (
        1 + lam * dt * char_func
    )


# ---------------------------------------------------------------------------
# A6  smooth_Heaviside.py:5-14
# ---------------------------------------------------------------------------


def smooth_Heaviside(H, phi, blend_w):
    """smooth_Heaviside.py:5-14"""
    H[...] = 0 * H
    H[...] = H + (phi >= blend_w)
    H[...] = H + (np.fabs(phi) < blend_w) * 0.5 * (
        1 + phi / blend_w + np.sin(np.pi * phi / blend_w) / np.pi
    )


# ---------------------------------------------------------------------------
# A7 / A13  ENO3 advection (pyst_kernels/advection_flux.py,
#           pyst_kernels/advection_timestep.py, pyst_kernels/elementwise_ops.py)
# ---------------------------------------------------------------------------

# pystencils infers the ghost-layer count from the largest stencil offset
# (2 for every ENO3 sub-kernel) and applies it to both dimensions, so the four
# flux loops cover rows/cols [2, n-2) of the (mirrored) grid.
ENO3_GHOST = 2


def _eno3_flux_conservative(field, velocity, inv_dx):
    """advection_flux.py:9-163.  Returns the flux array (zero on ghosts)."""
    g = ENO3_GHOST
    n0, n1 = field.shape
    flux = np.zeros_like(field)  # elementwise_ops.py:55-81 (set_fixed_val 0)
    f = field
    ux = velocity[0]
    uy = velocity[1]
    I = slice(g, n0 - g)
    J = slice(g, n1 - g)

    def sx(o):  # x (axis 1) offset view over the iteration space
        return (I, slice(g + o, n1 - g + o))

    def sy(o):  # y (axis 0) offset view over the iteration space
        return (slice(g + o, n0 - g + o), J)

    def fu(v, a, b, c):  # (1/3) f_a u_a + (5/6) f_b u_b - (1/6) f_c u_c
        return (
            (1 / 3) * f[a] * v[a] + (5 / 6) * f[b] * v[b] - (1 / 6) * f[c] * v[c]
        )

    # x front, advection_flux.py:32-44
    c = ux[sx(0)] > -ux[sx(1)]
    flux[I, J] = flux[I, J] + inv_dx * np.where(
        c, fu(ux, sx(1), sx(0), sx(-1)), fu(ux, sx(0), sx(1), sx(2))
    )
    # x back, advection_flux.py:61-73
    c = ux[sx(0)] > -ux[sx(-1)]
    flux[I, J] = flux[I, J] - inv_dx * np.where(
        c, fu(ux, sx(0), sx(-1), sx(-2)), fu(ux, sx(-1), sx(0), sx(1))
    )
    # y front, advection_flux.py:88-100
    c = uy[sy(0)] > -uy[sy(1)]
    flux[I, J] = flux[I, J] + inv_dx * np.where(
        c, fu(uy, sy(1), sy(0), sy(-1)), fu(uy, sy(0), sy(1), sy(2))
    )
    # y back, advection_flux.py:115-127
    c = uy[sy(0)] > -uy[sy(-1)]
    flux[I, J] = flux[I, J] - inv_dx * np.where(
        c, fu(uy, sy(0), sy(-1), sy(-2)), fu(uy, sy(-1), sy(0), sy(1))
    )
    return flux


def _eno3_flux_non_conservative(field, velocity, inv_dx):
    """advection_flux.py:166-332 (face interpolant times the cell velocity)."""
    g = ENO3_GHOST
    n0, n1 = field.shape
    flux = np.zeros_like(field)
    ux = velocity[0]
    uy = velocity[1]
    I = slice(g, n0 - g)
    J = slice(g, n1 - g)

    def sx(o):
        return slice(g + o, n1 - g + o)

    def sy(o):
        return slice(g + o, n0 - g + o)

    f = field
    # x front, advection_flux.py:189-202
    c = ux[I, sx(0)] > -ux[I, sx(1)]
    face = np.where(
        c,
        (1 / 3) * f[I, sx(1)] + (5 / 6) * f[I, sx(0)] - (1 / 6) * f[I, sx(-1)],
        (1 / 3) * f[I, sx(0)] + (5 / 6) * f[I, sx(1)] - (1 / 6) * f[I, sx(2)],
    )
    flux[I, J] = flux[I, J] + inv_dx * face * ux[I, J]
    # x back, advection_flux.py:219-236
    c = ux[I, sx(0)] > -ux[I, sx(-1)]
    face = np.where(
        c,
        (1 / 3) * f[I, sx(0)] + (5 / 6) * f[I, sx(-1)] - (1 / 6) * f[I, sx(-2)],
        (1 / 3) * f[I, sx(-1)] + (5 / 6) * f[I, sx(0)] - (1 / 6) * f[I, sx(1)],
    )
    flux[I, J] = flux[I, J] - inv_dx * face * ux[I, J]
    # y front, advection_flux.py:251-264
    c = uy[sy(0), J] > -uy[sy(1), J]
    face = np.where(
        c,
        (1 / 3) * f[sy(1), J] + (5 / 6) * f[sy(0), J] - (1 / 6) * f[sy(-1), J],
        (1 / 3) * f[sy(0), J] + (5 / 6) * f[sy(1), J] - (1 / 6) * f[sy(2), J],
    )
    flux[I, J] = flux[I, J] + inv_dx * face * uy[I, J]
    # y back, advection_flux.py:279-296
    c = uy[sy(0), J] > -uy[sy(-1), J]
    face = np.where(
        c,
        (1 / 3) * f[sy(0), J] + (5 / 6) * f[sy(-1), J] - (1 / 6) * f[sy(-2), J],
        (1 / 3) * f[sy(-1), J] + (5 / 6) * f[sy(0), J] - (1 / 6) * f[sy(1), J],
    )
    flux[I, J] = flux[I, J] - inv_dx * face * uy[I, J]
    return flux


def eno3_timestep_conservative(field, velocity, dt_by_dx):
    """advection_timestep.py:37-54: flux=0, 4 flux loops, field += flux."""
    flux = _eno3_flux_conservative(field, velocity, -dt_by_dx)
    field[...] = field + flux


def eno3_timestep_non_conservative(field, velocity, dt_by_dx):
    """advection_timestep.py:85-102"""
    flux = _eno3_flux_non_conservative(field, velocity, -dt_by_dx)
    field[...] = field + flux


def _mirror_velocity(u_z, u_r):
    """advect_vorticity_via_eno3.py:28-31: u_z even, u_r odd about the axis."""
    nr, nz = u_z.shape
    u_double = np.zeros((2, 2 * nr, nz))
    u_double[0, nr:, :] = u_z
    u_double[0, :nr, :] = np.flip(u_z, axis=0)
    u_double[1, nr:, :] = u_r
    u_double[1, :nr, :] = -np.flip(u_r, axis=0)
    return u_double


def advect_vorticity_via_eno3(vorticity, u_z, u_r, dt, dx):
    """advect_vorticity_via_eno3.py:20-42 (the closure body; omega odd)."""
    nr, nz = vorticity.shape
    u_double = _mirror_velocity(u_z, u_r)
    vd = np.zeros((2 * nr, nz))
    vd[nr:, :] = vorticity
    vd[:nr, :] = -np.flip(vorticity, axis=0)
    eno3_timestep_conservative(vd, u_double, dt / dx)
    vorticity[...] = vd[nr:, :]


def advect_refmap_via_eno3(eta1, eta2, u_z, u_r, dt, dx):
    """advect_refmap_via_eno3.py:20-52 (eta1 even, eta2 odd)."""
    nr, nz = eta1.shape
    u_double = _mirror_velocity(u_z, u_r)
    rd = np.zeros((2 * nr, nz))
    rd[nr:, :] = eta1
    rd[:nr, :] = np.flip(eta1, axis=0)
    eno3_timestep_non_conservative(rd, u_double, dt / dx)
    eta1[...] = rd[nr:, :]
    rd[nr:, :] = eta2
    rd[:nr, :] = -np.flip(eta2, axis=0)
    eno3_timestep_non_conservative(rd, u_double, dt / dx)
    eta2[...] = rd[nr:, :]


# ---------------------------------------------------------------------------
# A8  diffusion_RK2.py:4-44
# ---------------------------------------------------------------------------


def diffusion_RK2_unb(vorticity, temp_vorticity, R, nu, dt, dx):
    """diffusion_RK2.py:4-44"""
    _rk2_stage1(vorticity, temp_vorticity, R, nu, dt, dx)
    _rk2_stage2(vorticity, temp_vorticity, R, nu, dt, dx)


def _rk2_stage1(vorticity, temp_vorticity, R, nu, dt, dx):
    """diffusion_RK2.py:9-26 (and :52-70 of the periodic form)"""
    temp_vorticity[...] = vorticity
    temp_vorticity[1:-1, 1:-1] += (
        0.5
        * nu
        * dt
        * (
            (
                vorticity[2:, 1:-1]
                + vorticity[:-2, 1:-1]
                + vorticity[1:-1, 2:]
                + vorticity[1:-1, :-2]
                - 4 * vorticity[1:-1, 1:-1]
            )
            / (dx**2)
            + (vorticity[2:, 1:-1] - vorticity[:-2, 1:-1]) / (2 * dx) / R[1:-1, 1:-1]
            - vorticity[1:-1, 1:-1] * (R[1:-1, 1:-1] ** -2)
        )
    )


def _rk2_stage2(vorticity, temp_vorticity, R, nu, dt, dx):
    """diffusion_RK2.py:27-44 (and :72-90 of the periodic form)"""
    vorticity[1:-1, 1:-1] += (
        nu
        * dt
        * (
            (
                temp_vorticity[2:, 1:-1]
                + temp_vorticity[:-2, 1:-1]
                + temp_vorticity[1:-1, 2:]
                + temp_vorticity[1:-1, :-2]
                - 4 * temp_vorticity[1:-1, 1:-1]
            )
            / (dx**2)
            + (temp_vorticity[2:, 1:-1] - temp_vorticity[:-2, 1:-1])
            / (2 * dx)
            / R[1:-1, 1:-1]
            - temp_vorticity[1:-1, 1:-1] * (R[1:-1, 1:-1] ** -2)
        )
    )


# ---------------------------------------------------------------------------
# A9  compute_forces.py:4-17 and the inline drag coefficient
#     examples/FlowPastSphere/flow_past_sphere.py:166-175
# ---------------------------------------------------------------------------


def compute_force_on_body(
    R, part_char_func, rho_f, brink_lam, u_z, U_z_cm_part, part_vol, dt, diff
):
    """compute_forces.py:4-17"""
    F_pen = rho_f * brink_lam * np.sum(R * part_char_func * (u_z - U_z_cm_part))
    F_un = (diff * part_vol) / dt
    return F_pen, F_un


def drag_coefficient(R, char_func, u_z, dx, brink_lam, r_sph):
    """flow_past_sphere.py:166-175 (uses the penalised u_z)."""
    return (
        2 * 2 * np.pi * dx * dx * brink_lam * np.sum(R * char_func * u_z)
        / (np.pi * r_sph**2)
    )


# ---------------------------------------------------------------------------
# A10 dt rule, flow_past_sphere.py:150-153
# ---------------------------------------------------------------------------


def max_abs_velocity_sum(u_z, u_r):
    """np.amax(np.fabs(u_z) + np.fabs(u_r)) as in flow_past_sphere.py:152"""
    return np.amax(np.fabs(u_z) + np.fabs(u_r))


def cfl_dt(u_z, u_r, dx, nu, CFL=0.1):
    """flow_past_sphere.py:150-153"""
    eps = np.finfo(float).eps
    return min(
        0.9 * dx**2 / 4 / nu,
        CFL * dx / (np.amax(np.fabs(u_z) + np.fabs(u_r)) + eps),
    )


# ---------------------------------------------------------------------------
# A11 solid_sigma.py:4-29
# ---------------------------------------------------------------------------


def solid_sigma(
    sigma_s_11, sigma_s_12, sigma_s_22, G, dx, eta_1, eta_2, eta_1z, eta_1r, eta_2z,
    eta_2r,
):
    """solid_sigma.py:4-29 (unwritten scratch entries keep their values)."""
    eta_1z[:, 1:-1] = (eta_1[:, 2:] - eta_1[:, :-2]) / (2 * dx)
    eta_2z[:, 1:-1] = (eta_2[:, 2:] - eta_2[:, :-2]) / (2 * dx)
    eta_1r[1:-1, 1:-1] = (eta_1[2:, 1:-1] - eta_1[:-2, 1:-1]) / (2 * dx)
    eta_1r[0, :] = (-eta_1[2, :] + 4 * eta_1[1, :] - 3 * eta_1[0, :]) / (2 * dx)
    eta_2r[1:-1, 1:-1] = (eta_2[2:, 1:-1] - eta_2[:-2, 1:-1]) / (2 * dx)
    eta_2r[0, :] = (-eta_2[2, :] + 4 * eta_2[1, :] - 3 * eta_2[0, :]) / (2 * dx)
    sigma_s_12[...] = -G * (eta_1z * eta_1r + eta_2z * eta_2r)
    sigma_s_11[...] = (0.5 * G) * (eta_1r**2 + eta_2r**2 - eta_1z**2 - eta_2z**2)
    sigma_s_22[...] = -sigma_s_11


# ---------------------------------------------------------------------------
# A12 div_tau.py:4-34
# ---------------------------------------------------------------------------


def update_vorticity_from_solid_stress(
    vorticity, tau_z, tau_r, tau11, tau12, tau22, R, dt, dx
):
    """div_tau.py:4-34"""
    _div_tau(tau_z, tau_r, tau11, tau12, tau22, R, dx)
    _curl_tau(vorticity, tau_z, tau_r, dt, dx)


def _div_tau(tau_z, tau_r, tau11, tau12, tau22, R, dx):
    """div_tau.py:7-28 (and :45-67 of the periodic form)"""
    tau_z[1:-1, 1:-1] = (
        tau11[1:-1, 2:] - tau11[1:-1, :-2] + tau12[2:, 1:-1] - tau12[:-2, 1:-1]
    ) / (2 * dx) + tau12[1:-1, 1:-1] / R[1:-1, 1:-1]
    tau_z[0, 1:-1] = (
        tau11[0, 2:]
        - tau11[0, :-2]
        - tau12[2, 1:-1]
        + 4 * tau12[1, 1:-1]
        - 3 * tau12[0, 1:-1]
    ) / (2 * dx) + tau12[0, 1:-1] / R[0, 1:-1]
    tau_r[1:-1, 1:-1] = (
        tau12[1:-1, 2:] - tau12[1:-1, :-2] + tau22[2:, 1:-1] - tau22[:-2, 1:-1]
    ) / (2 * dx) + tau22[1:-1, 1:-1] / R[1:-1, 1:-1]
    tau_r[0, 1:-1] = (
        tau12[0, 2:]
        - tau12[0, :-2]
        - tau22[2, 1:-1]
        + 4 * tau22[1, 1:-1]
        - 3 * tau22[0, 1:-1]
    ) / (2 * dx) + tau22[0, 1:-1] / R[0, 1:-1]


def _curl_tau(vorticity, tau_z, tau_r, dt, dx):
    """div_tau.py:29-34 (and :71-75 of the periodic form)"""
    vorticity[1:-1, 1:-1] += (
        dt
        * (tau_r[1:-1, 2:] - tau_r[1:-1, :-2] - tau_z[2:, 1:-1] + tau_z[:-2, 1:-1])
        / (2 * dx)
    )


# ---------------------------------------------------------------------------
# A14 host glue: extrapolate_eta_using_least_squares_unb.py:7-30 and
#     extrapolate_using_least_squares.py:8-38 (the sweep itself is lsq.c)
# ---------------------------------------------------------------------------


def extrapolate_eta_with_least_squares(
    inside_solid, ball_phi, eta1, eta2, ball_phi_double, eta1_double, eta2_double,
    extrap_zone, grid_size_r, z,
):
    """extrapolate_eta_using_least_squares_unb.py:7-30"""
    from oracle.lsq import extrapolate_using_least_squares_till_first_order

    ball_phi_double[grid_size_r:, :] = -ball_phi
    eta1_double[grid_size_r:, :] = inside_solid * eta1
    eta2_double[grid_size_r:, :] = inside_solid * eta2
    ball_phi_double[:grid_size_r, :] = -np.flip(ball_phi, axis=0)
    eta1_double[:grid_size_r, :] = np.flip(eta1_double[grid_size_r:, :], axis=0)
    eta2_double[:grid_size_r, :] = -np.flip(eta2_double[grid_size_r:, :], axis=0)
    # extrapolate_using_least_squares.py:33-37
    current_flag = np.asarray(ball_phi_double < 0, dtype=np.int16)
    target_flag = np.asarray(ball_phi_double < extrap_zone, dtype=np.int16)
    extrapolate_using_least_squares_till_first_order(
        current_flag, target_flag, eta1_double, eta2_double, z, z
    )
    eta1[...] = eta1_double[grid_size_r:, :]
    eta2[...] = eta2_double[grid_size_r:, :]
