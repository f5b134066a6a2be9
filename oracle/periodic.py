"""Periodic-z kernels (test oracle only): the ghost exchange of
kernels/periodic_boundary_ghost_comm.py:4-32 and the *_periodic forms of
the stencils, each the reference's exchange calls around the unbounded
bodies restated in oracle.kernels:
compute_velocity_from_psi.py:20-35, compute_vorticity_from_velocity.py:16-28,
diffusion_RK2.py:47-90, advect_vorticity_via_eno3.py:47-91,
advect_refmap_via_eno3.py:57-116, solid_sigma.py:32-60, div_tau.py:37-75.
"""
from oracle import kernels as K


def gen_periodic_boundary_ghost_comm(ghost_size):
    """:4-15"""
    g = ghost_size

    def comm(field):
        field[:, :g] = field[:, -2 * g:-g]
        field[:, -g:] = field[:, g:2 * g]

    return comm


def gen_periodic_boundary_ghost_comm_eta(ghost_size, Z_max, dx):
    """:18-32"""
    g = ghost_size

    def comm(field):
        field[:, :g] = field[:, -2 * g:-g] - Z_max + 2 * g * dx
        field[:, -g:] = field[:, g:2 * g] + Z_max - 2 * g * dx

    return comm


def compute_velocity_from_psi_periodic(u_z, u_r, psi, R, dx, per_communicator):
    per_communicator(psi)
    K.compute_velocity_from_psi_unb(u_z, u_r, psi, R, dx)


def compute_vorticity_from_velocity_periodic(vort, u_z, u_r, dx, per_communicator):
    per_communicator(u_r)
    per_communicator(u_z)
    K.compute_vorticity_from_velocity_unb(vort, u_z, u_r, dx)


def diffusion_RK2_periodic(vorticity, temp_vorticity, R, nu, dt, dx, per_communicator):
    per_communicator(vorticity)
    K._rk2_stage1(vorticity, temp_vorticity, R, nu, dt, dx)
    per_communicator(temp_vorticity)
    K._rk2_stage2(vorticity, temp_vorticity, R, nu, dt, dx)


def advect_vorticity_via_eno3_periodic(vorticity, u_z, u_r, dt, dx, per_communicator):
    per_communicator(u_z)
    per_communicator(u_r)
    per_communicator(vorticity)
    K.advect_vorticity_via_eno3(vorticity, u_z, u_r, dt, dx)


def advect_refmap_via_eno3_periodic(eta1, eta2, u_z, u_r, dt, dx, per_communicator1,
                                    per_communicator2):
    per_communicator1(u_z)
    per_communicator1(u_r)
    per_communicator2(eta1)
    per_communicator1(eta2)
    K.advect_refmap_via_eno3(eta1, eta2, u_z, u_r, dt, dx)


def solid_sigma_periodic(sigma_s_11, sigma_s_12, sigma_s_22, G, dx, eta_1, eta_2, eta_1z,
                         eta_1r, eta_2z, eta_2r, per_communicator1, per_communicator2):
    per_communicator2(eta_1)
    per_communicator1(eta_2)
    K.solid_sigma(sigma_s_11, sigma_s_12, sigma_s_22, G, dx, eta_1, eta_2, eta_1z, eta_1r,
                  eta_2z, eta_2r)


def update_vorticity_from_solid_stress_periodic(vorticity, tau_z, tau_r, tau11, tau12, tau22,
                                                R, dt, dx, per_communicator):
    per_communicator(tau11)
    per_communicator(tau12)
    per_communicator(tau22)
    K._div_tau(tau_z, tau_r, tau11, tau12, tau22, R, dx)
    per_communicator(tau_r)
    per_communicator(tau_z)
    K._curl_tau(vorticity, tau_z, tau_r, dt, dx)
