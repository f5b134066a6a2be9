/*
 * axisym.h -- C ABI of libaxisym_hip.so, the MI355X (gfx950) kernels for
 * PyAxisymFlow's per-timestep vorticity-streamfunction path on the
 * axisymmetric (z, r) grid.
 *
 * Conventions
 *   - Fields are row-major (nr, nz) arrays in device memory: row i = radial
 *     index (r), column j = axial index (z, contiguous).  "n" is nr*nz.
 *   - Every pointer is a device pointer unless the name ends in _host.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Calls are stream-ordered and asynchronous unless stated otherwise.
 *   - Every function returns axs_status; on failure axs_last_error() holds
 *     a message (thread-local).  Arguments are validated before any launch.
 *   - Scalars that the reference computes on the host (dt) are passed by
 *     value; the *_dev variants used by the fused stepper read them from
 *     device memory instead.
 * Citations name the reference interface each entry point replaces
 * (paths inside the PyAxisymFlow tree).
 */
#ifndef AXISYM_H_
#define AXISYM_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int axs_status;
#define AXS_OK 0
#define AXS_ERR_INVALID 1
#define AXS_ERR_HIP 2
#define AXS_ERR_UNSUPPORTED 3
#define AXS_ERR_ALLOC 4

/* Thread-local message for the last failing call on this thread. */
const char* axs_last_error(void);
/* Library version string. */
const char* axs_version(void);

/* ---- A1: kernels/kill_boundary_vorticity_sine.py:4-14 (z) and :17-27 (r).
 * Z and R are the full (nr, nz) coordinate arrays the reference receives.
 * Requires nz >= width, nr >= width, width >= 2. */
axs_status axs_kill_boundary_vorticity_sine_z_f64(double* vorticity, const double* Z,
                                                  int nr, int nz, int width, double dx,
                                                  void* stream);
axs_status axs_kill_boundary_vorticity_sine_r_f64(double* vorticity, const double* R,
                                                  int nr, int nz, int width, double dx,
                                                  void* stream);

/* ---- A3: kernels/compute_velocity_from_psi.py:4-17 (nr, nz >= 3). */
axs_status axs_compute_velocity_from_psi_unb_f64(double* u_z, double* u_r,
                                                 const double* psi, const double* R,
                                                 int nr, int nz, double dx, void* stream);

/* ---- A4: kernels/compute_vorticity_from_velocity.py:4-13 (interior only). */
axs_status axs_compute_vorticity_from_velocity_unb_f64(double* vort, const double* u_z,
                                                       const double* u_r, int nr, int nz,
                                                       double dx, void* stream);

/* ---- A5: kernels/brinkmann_penalize.py:4-16 (elementwise over n cells). */
axs_status axs_brinkmann_penalize_f64(double lam, double dt, const double* char_func,
                                      double U_z, double U_r, const double* grid_u_z,
                                      const double* grid_u_r, double* penalized_u_z,
                                      double* penalized_u_r, int64_t n, void* stream);

/* ---- A6: kernels/smooth_Heaviside.py:5-14 (elementwise over n cells). */
axs_status axs_smooth_heaviside_f64(double* H, const double* phi, double blend_w,
                                    int64_t n, void* stream);

/* ---- A7: kernels/advect_vorticity_via_eno3.py:8-44 with
 * pyst_kernels/advection_timestep.py:14-56 and advection_flux.py:9-163.
 * One conservative ENO3 Euler step of omega (odd about the axis), mirrored
 * in-index (no doubled buffers).  `scratch` is an (nr, nz) device buffer.
 * Requires nr >= 3, nz >= 5. */
axs_status axs_advect_vorticity_eno3_f64(double* vorticity, const double* u_z,
                                         const double* u_r, double* scratch, int nr,
                                         int nz, double dt, double dx, void* stream);

/* ---- A8: kernels/diffusion_RK2.py:4-44 (temp_vorticity is in/out scratch). */
axs_status axs_diffusion_rk2_unb_f64(double* vorticity, double* temp_vorticity,
                                     const double* R, int nr, int nz, double nu,
                                     double dt, double dx, void* stream);

/* ---- A9: kernels/compute_forces.py:4-17 (F_pen's sum) and the drag
 * coefficient sum of examples/FlowPastSphere/flow_past_sphere.py:166-175.
 * Writes sum(R * chi * (u_z - U)) to out_sum[0] (device, FP64 always).
 * `partials` is device scratch of axs_reduce_partials_size(n) doubles. */
int64_t axs_reduce_partials_size(int64_t n);
axs_status axs_penalization_force_sum_f64(const double* R, const double* chi,
                                          const double* u_z, double U, int64_t n,
                                          double* partials, double* out_sum,
                                          void* stream);

/* Allocation-free variant of axs_penalization_force_sum_f64: no counterpart
 * in the reference; for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same result, bit for bit; the partial sums live in a
 * library-owned device array of 1024 doubles (8 KiB per GPU, held until
 * the process ends) instead of a caller workspace, and the call allocates
 * nothing.  n <= AXS_NOALLOC_MAX_CELLS.  Calls on one GPU share that
 * array: calls that overlap on one GPU (different streams, or threads)
 * race and must be serialised by the caller. */
#define AXS_NOALLOC_MAX_CELLS ((int64_t)1 << 40)
axs_status axs_penalization_force_sum_noalloc_f64(const double* R, const double* chi,
                                                  const double* u_z, double U, int64_t n,
                                                  double* out_sum, void* stream);

/* ---- A10: the dt rule's reduction np.amax(np.fabs(u_z) + np.fabs(u_r))
 * (examples/FlowPastSphere/flow_past_sphere.py:150-153).  Writes to
 * out_max[0] (device). */
axs_status axs_max_abs_velocity_sum_f64(const double* u_z, const double* u_r, int64_t n,
                                        double* partials, double* out_max, void* stream);

/* Plain device copy of n doubles (16-byte aligned): no counterpart in the
 * reference; the achievable-bandwidth reference and the calibration kernel
 * for hardware byte counters. */
axs_status axs_device_copy_f64(double* dst, const double* src, int64_t n, void* stream);

/* ---- A11: elasto_kernels/solid_sigma.py:4-29.  eta_1z..eta_2r are in/out
 * scratch with the reference's partial-write semantics. */
axs_status axs_solid_sigma_f64(double* sigma_s_11, double* sigma_s_12, double* sigma_s_22,
                               double G, double dx, const double* eta_1,
                               const double* eta_2, double* eta_1z, double* eta_1r,
                               double* eta_2z, double* eta_2r, int nr, int nz,
                               void* stream);

/* ---- A12: elasto_kernels/div_tau.py:4-34 (tau_z, tau_r in/out scratch). */
axs_status axs_update_vorticity_from_solid_stress_f64(double* vorticity, double* tau_z,
                                                      double* tau_r, const double* tau11,
                                                      const double* tau12,
                                                      const double* tau22, const double* R,
                                                      int nr, int nz, double dt, double dx,
                                                      void* stream);

/* ---- Periodic z: kernels/periodic_boundary_ghost_comm.py:4-32.  Ghost
 * columns of a z-periodic field: field[:, :g] = field[:, -2g:-g],
 * field[:, -g:] = field[:, g:2g]; eta_shift != 0 is the reference map's
 * version (gen_periodic_boundary_ghost_comm_eta: the left ghosts get
 * - z_max + 2 g dx, the right ones + z_max - 2 g dx).  Needs nz >= 4 g. */
axs_status axs_periodic_ghost_comm_f64(double* field, int nr, int nz, int ghost, int eta_shift,
                                       double z_max, double dx, void* stream);
/* kernels/diffusion_RK2.py:47-90 (diffusion_RK2_periodic): the ghost
 * exchange of `ghost` columns before each RK2 stage. */
axs_status axs_diffusion_rk2_periodic_f64(double* vorticity, double* temp_vorticity,
                                          const double* R, int nr, int nz, double nu, double dt,
                                          double dx, int ghost, void* stream);
/* elasto_kernels/div_tau.py:37-75 (update_vorticity_from_solid_stress_periodic):
 * ghost exchange of tau11, tau12, tau22 (in place), tau_z and tau_r, then
 * of tau_r and tau_z before the vorticity update. */
axs_status axs_update_vorticity_from_solid_stress_periodic_f64(
    double* vorticity, double* tau_z, double* tau_r, double* tau11, double* tau12, double* tau22,
    const double* R, int nr, int nz, double dt, double dx, int ghost, void* stream);

/* ---- A13: elasto_kernels/advect_refmap_via_eno3.py:8-54 with the
 * non-conservative flux of pyst_kernels/advection_flux.py:166-332.
 * eta1 is even and eta2 odd about the axis; `scratch` holds 2 (nr, nz) buffers. */
axs_status axs_advect_refmap_eno3_f64(double* eta1, double* eta2, const double* u_z,
                                      const double* u_r, double* scratch, int nr, int nz,
                                      double dt, double dx, void* stream);

/* ---- A2: kernels/FastDiagonalisationStokesSolver.py:6-156.
 * Solves A_r psi + psi P_z^T = r * rhs with P_z, P_r, D_r as built at
 * :41-97 for each bc_type (Neumann z and r; Neumann or Dirichlet r with
 * periodic z), by a DCT-II (periodic: a real DFT) along z and one radial
 * tridiagonal solve per z-mode.  The plan owns its device tables and
 * scratch. */
typedef struct axs_stokes_plan axs_stokes_plan;
#define AXS_BC_NEUMANN_Z_AND_R 0
/* :88-97: periodic P_z (real DFT along z); Neumann or Dirichlet at r_max */
#define AXS_BC_NEUMANN_R_PERIODIC_Z 1
#define AXS_BC_DIRICHLET_R_PERIODIC_Z 2
axs_status axs_stokes_plan_create(int nr, int nz, double dx, int bc_type,
                                  axs_stokes_plan** plan_out);
axs_status axs_stokes_plan_destroy(axs_stokes_plan* plan);
/* psi and rhs may alias. */
axs_status axs_stokes_solve_f64(axs_stokes_plan* plan, double* psi, const double* rhs,
                                void* stream);

/* ---- Velocity potential: kernels/FastDiagonalisationPotentialSolver.py:6-145
 * (called by torus_streaming.py:176).  Solves A_r phi + phi P_z^T = rhs with
 * P_z, P_r, D_r as built at :41-86 (Neumann z and r), by a DCT-II along z
 * and one radial tridiagonal solve per z-mode.  Mode 0's radial system is
 * singular (constants are in the null space); it is solved in the gauge
 * phi_hat_0[nr - 1] = 0, where the reference divides by a round-off
 * eigenvalue (:114-117).  phi therefore matches the reference up to an
 * additive constant when rhs is compatible, and its gradient does not see
 * the constant.  phi and rhs may alias. */
typedef struct axs_potential_plan axs_potential_plan;
axs_status axs_potential_plan_create(int nr, int nz, double dx, int bc_type,
                                     axs_potential_plan** plan_out);
axs_status axs_potential_plan_destroy(axs_potential_plan* plan);
axs_status axs_potential_solve_f64(axs_potential_plan* plan, double* phi, const double* rhs,
                                   void* stream);
/* kernels/compute_velocity_from_phi.py:4-17 (compute_velocity_from_phi_unb):
 * u_z = d phi / dz, u_r = d phi / dr, second-order one-sided at the edges. */
axs_status axs_velocity_from_phi_f64(double* u_z, double* u_r, const double* phi, int nr, int nz,
                                     double dx, void* stream);

/* ---- Implicit Euler diffusion: kernels/implicit_diffusion_solver.py:6-143
 * (ImplicitEulerDiffusionStepper; called by flow_past_sphere.py:180-181,
 * spheroid_streaming.py:236-237, torus_streaming.py:182 when
 * implicit_diffusion is set).  Solves (I - nu_dt L) w' = w in place, with
 * L w = A_r w + w P_z^T built as :40-72 (zero ghost values at every edge:
 * Dirichlet P_z, A_r = P_r + D_r / r - I / r^2), by a DST-I along z, one
 * radial tridiagonal solve per z-mode and the DST-I back.  nu_dt is
 * kinematic_viscosity * time_step, fixed per plan (the reference rejects any
 * other dt, :115-120).  The plan owns its device tables and scratch. */
typedef struct axs_implicit_plan axs_implicit_plan;
axs_status axs_implicit_diffusion_plan_create(int nr, int nz, double dx, double nu_dt,
                                              axs_implicit_plan** plan_out);
axs_status axs_implicit_diffusion_plan_destroy(axs_implicit_plan* plan);
axs_status axs_implicit_diffusion_step_f64(axs_implicit_plan* plan, double* omega,
                                           void* stream);

/* ---- A14: core/src/extrapolate_using_least_squares.hpp:450-467
 * (extrapolate_using_least_squares_till_first_order<short, double>, bound
 * by core/extrapolate_using_least_squares_bind.cpp:11-41).  Flags are int16
 * (n0, n1); grid_x has n1 entries and grid_y n0.  Sweeps until no cell is
 * extrapolated or max_sweeps is reached; sweeps_done_host (may be NULL)
 * receives the count.  Synchronizes `stream` (the sweep count is data
 * dependent).  Out-of-range window cells read as unknown (flag 0). */
axs_status axs_extrapolate_lsq_first_order_f64(int16_t* current_flag,
                                               const int16_t* target_flag, double* eta_x,
                                               double* eta_y, const double* grid_x,
                                               const double* grid_y, int n0, int n1,
                                               int max_sweeps, int* sweeps_done_host,
                                               void* stream);

/* ---- A15: narrow-band signed-distance reinitialisation.  Replaces the
 * call skfmm.distance(phi, dx=dx, narrow=band) (scikit-fmm 2022.8.15,
 * order 2) at examples/SoftSphereInTaylorGreenVortex/
 * soft_sphere_in_taylor_green_vortex.py:159-162.  out = signed distance to
 * phi's zero level set where |d| <= narrow and phi elsewhere; masked[c] = 1
 * (if masked is not null) where out keeps phi, as skfmm's masked array.
 * The discrete equations (linear interface distances, second-order upwind
 * eikonal update) are solved by `iterations` Jacobi sweeps instead of fast
 * marching (< 0: 2 ceil(narrow / dx) + 8).  phi and out may alias. */
axs_status axs_reinit_distance_f64(const double* phi, double* out, uint8_t* masked, int nr,
                                   int nz, double dx, double narrow, int iterations,
                                   void* stream);

/* ---- Particle remeshing with the M'4 kernel (next to the path: the
 * alternative advection of kernels/advect_particle.py:5-68, used by the
 * pipe-flow and particle-in-bubble drivers).
 * P2M: core/src/particles_to_mesh.hpp:71-91 (periodic) and :165-185
 * (unbounded), bound as particles_to_mesh_2D_mp4 / _unbounded_mp4
 * (core/particles_to_mesh_bind.cpp:35-90): q at particles (px, py) onto the
 * (n0, n1) mesh `out` (cleared first), cell centres at (k + 1/2) d.  The
 * particle arrays have the mesh's shape (n0, n1): each cell gathers from the
 * particles within D + 2 of it, D the largest distance (in cells) of a
 * particle's nearest mesh index from its array index, measured first
 * (synchronizes `stream`; max_displacement_host, may be NULL, receives D).
 * Results equal the reference's scatter bit for bit.  D > 4 (or a periodic
 * mesh below max(2 D + 4, 8) cells along an axis) returns
 * AXS_ERR_UNSUPPORTED. */
axs_status axs_particles_to_mesh_2d_mp4_f64(const double* px, const double* py, const double* q,
                                            double* out, int n0, int n1, double dx, double dy,
                                            int periodic, int* max_displacement_host,
                                            void* stream);
/* M2P: core/src/mesh_to_particles.hpp:111-131 (periodic) and :209-229
 * (unbounded), bound as mesh_to_particles_2D_mp4 / _unbounded_mp4: the
 * (m0, m1) mesh fields fx, fy at np particles (px, py) into ox, oy.  Any
 * particle count and layout.  A particle too far out for a mesh index
 * (|x / dx| or |y / dy| >= 1e9, where the reference is undefined) gets NaN. */
axs_status axs_mesh_to_particles_2d_mp4_f64(const double* fx, const double* fy, const double* px,
                                            const double* py, double* ox, double* oy, int m0,
                                            int m1, int64_t np, double dx, double dy,
                                            int periodic, void* stream);
/* kernels/advect_particle.py:5-35 (unbounded) and :38-68 (periodic): the
 * particles of the doubled (2 nr, nz) grid pushed by the mirrored velocity
 * (u_z even, u_r odd about the axis), the mirrored vorticity (odd) remeshed
 * by P2M into vort_double (a (2 nr, nz) work buffer), the particles reset to
 * (Z_double, R_double) and vorticity = the upper half.  Synchronizes `stream`. */
axs_status axs_advect_vorticity_via_particles_f64(double* z_particles, double* r_particles,
                                                  double* vort_particles, double* vorticity,
                                                  const double* Z_double, const double* R_double,
                                                  const double* u_z, const double* u_r,
                                                  double* vort_double, int nr, int nz, double dx,
                                                  double dt, int periodic, void* stream);

/* ---- A16: the rigid-body driver step, device resident.
 * examples/FlowPastSphere/flow_past_sphere.py:107-188 (static body, free
 * stream with the sine ramp, drag coefficient Cd) and
 * examples/SpheroidStreaming/spheroid_streaming.py:129-246 (oscillating
 * spheroid: mask rebuilt each step, penalisation velocity U_0 cos(omega t),
 * dt clipped to the period and tEnd, psi averaged per period).
 * One step: kill taper -> Stokes solve -> velocity (+ free stream) -> dt ->
 * [psi average] -> [body mask] -> Brinkman + curl -> force -> ENO3 -> RK2,
 * with t, dt and the force history on the device (scalar slots: t, dt,
 * dt/dx, lam dt, nu dt/2, nu dt, ramp factor, U_pen, max|u|, force sum,
 * step counter, period timer, average reset flag, body centre, period timer
 * of the step, ramp factor and U_pen of the last completed step).
 * work0 holds the other half of a ping-pong pair with `vorticity`; when
 * axs_rigid_stepper_run returns, `vorticity` holds the current field.
 * u_pen_z / u_pen_r are filled on demand by axs_rigid_stepper_velocity;
 * u_z, u_r and work1 are unused by the step (kept for the layout).
 * precision AXS_F64: every field FP64.  AXS_F32: fields, stencils and
 * z-transforms FP32; radial tridiagonal, mask level set, reductions and
 * step scalars FP64.  The caller owns every buffer: (nr, nz) fields of the
 * chosen precision, scalars (AXS_RIGID_SCALARS doubles), drag (history doubles or NULL),
 * r (nr doubles), z (nz doubles; moving body only), kill_wz (2*kill_width
 * doubles: left then right block), kill_wr (kill_width doubles). */
#define AXS_F64 0
#define AXS_F32 1
#define AXS_RIGID_SCALARS 24
typedef struct axs_rigid_stepper axs_rigid_stepper;
typedef struct {
  int nr, nz, kill_width, history;
  int precision;       /* AXS_F64 or AXS_F32 */
  int moving_body;     /* 1: chi = H(phi) of the oscillating spheroid, rebuilt each step */
  int average_psi;     /* 1: avg_psi += psi dt, restarted at each period boundary */
  int implicit_diffusion; /* 1: implicit Euler diffusion (implicit_diffusion_solver.py)
                          * replaces the RK2 stages, dt = dt_diffusive every step */
  double dx, nu, brink_lam;
  double U0;           /* free stream added to u_z each step */
  double T_ramp;       /* ramp sin(pi t / (2 T_ramp)) while t < T_ramp; <= 0: none */
  double CFL, eps;     /* dt = min(dt_diffusive, CFL dx / (max|u| + eps), dt_cap) */
  double dt_diffusive; /* 0.9 dx^2 / 4 / nu as the driver forms it; with
                        * implicit_diffusion the fixed dt (dx^2 / 4 / nu) */
  double osc_amp, osc_omega; /* penalisation velocity U_z(t) = osc_amp cos(osc_omega t) */
  double drag_scale, drag_div; /* force = drag_scale * sum(r chi (u_z - U_z(t))) / drag_div */
  double dt_cap;       /* third term of the dt min (HUGE_VAL: none) */
  double period;       /* clip dt to period boundaries (<= 0: none) */
  double t_end;        /* clip dt at t_end (<= 0: none) */
  /* moving spheroid: phi = radius - sqrt((z - zc(t))^2 / aspect2 + (r - r0)^2),
   * zc(t) = z0 + disp sin(osc_omega t), chi = smooth Heaviside(phi, moll_zone) */
  double body_z0, body_r0, body_aspect2, body_radius, body_disp, moll_zone;
} axs_rigid_config;
typedef struct {
  void *vorticity, *psi, *u_z, *u_r, *u_pen_z, *u_pen_r, *char_func, *work0, *work1, *avg_psi;
  double* scalars;
  double* drag;
  const double* r;
  const double* z;
  const double* kill_wz;
  const double* kill_wr;
} axs_rigid_buffers;
axs_status axs_rigid_stepper_create(const axs_rigid_config* cfg, const axs_rigid_buffers* bufs,
                                    axs_rigid_stepper** out);
/* The batched form: `batch` cases stepped together by the same launches
 * (the case index is a grid dimension of every kernel).  cfgs[batch] share
 * kill_width, history, precision and the moving_body / average_psi flags;
 * on one grid they share nr, nz and dx and may differ in everything else
 * (viscosity, body
 * mask, ramp, dt rule, force scale).  Per-case buffers hold the cases back
 * to back: fields batch * nr * nz elements, scalars batch *
 * AXS_RIGID_SCALARS, drag batch * history; r, z, kill_wz and kill_wr are
 * shared.  Each case advances its own t and dt.  The reference steps one
 * case per process (flow_past_sphere.py:107-188); this runs that loop body
 * for many cases at once.  axs_rigid_stepper_create is batch_create with
 * batch = 1.
 *
 * Uneven batches: cases may also differ in nr, nz and dx when none has a
 * moving body, every nz <= 1024, and the cases come in at most 16 runs of
 * one grid.  Every pass is then one launch over all cases.  Fields hold
 * case q's nr_q * nz_q elements after those of cases 0..q-1; r holds each
 * case's nr_q radii back to back, kill_wz each case's 2 kill_width
 * weights, kill_wr each case's kill_width weights (z is unused). */
axs_status axs_rigid_batch_create(const axs_rigid_config* cfgs, int batch,
                                  const axs_rigid_buffers* bufs, axs_rigid_stepper** out);
axs_status axs_rigid_stepper_destroy(axs_rigid_stepper* st);
/* Queues nsteps steps on `stream`; use_graph != 0 captures the step (one
 * graph per ping-pong direction) on first use and replays it (stream must
 * not be the null stream). */
axs_status axs_rigid_stepper_run(axs_rigid_stepper* st, int nsteps, int use_graph, void* stream);
/* Graph launches queued so far by axs_rigid_stepper_run with use_graph:
 * graphs of 16 steps ("chunks") and shorter graphs of 8, 4, 2 or 1 steps (a
 * run of n steps launches n / 16 chunks, from either ping-pong parity, then
 * one graph per set bit of n mod 16; its last graph also completes the
 * run).  Diagnostics; no counterpart in the reference. */
axs_status axs_rigid_stepper_graph_launches(const axs_rigid_stepper* st, int64_t* chunks,
                                            int64_t* singles);
/* Profiling hook: the step is a fixed sequence of stages (taper + dct2 rows,
 * tridiagonal columns, dct3 rows, velocity maximum, body mask, fused
 * penalisation / advection / diffusion, step scalars; disabled stages are
 * empty).  run_stage queues `reps` launches of one stage alone, so a timer
 * on `stream` measures that stage.  It advances no state consistently: use
 * it only on a stepper whose results are discarded. */
int axs_rigid_stepper_stage_count(void);
const char* axs_rigid_stepper_stage_name(int stage);
axs_status axs_rigid_stepper_run_stage(axs_rigid_stepper* st, int stage, int reps, void* stream);
/* Writes the penalised velocity of the last completed step (the velocity
 * ENO3 advected with: brinkmann_penalize.py:4-16 applied to
 * compute_velocity_from_psi.py:4-17 plus the free stream) into u_pen_z and
 * u_pen_r.  Not part of the step, so a caller that never reads the
 * velocity pays nothing for it. */
axs_status axs_rigid_stepper_velocity(axs_rigid_stepper* st, void* stream);
/* Recomputes the ramp factor and U_pen for the time now in scalars[0]. */
axs_status axs_rigid_stepper_reset_scalars(axs_rigid_stepper* st, void* stream);

/* ---- A16 (soft body): the soft-sphere driver step, device resident.
 * examples/SoftSphereInTaylorGreenVortex/soft_sphere_in_taylor_green_vortex.py:101-213:
 * kill taper -> Stokes solve -> velocity -> dt = min(CFL dx / (max|u| +
 * eps), CFL dx / sqrt(G / rho_f), 0.9 dx^2 / 4 / nu) -> reference-map ENO3
 * -> phi pinned to the sphere of eta where phi > -3 dx -> narrow-band
 * reinit (as axs_reinit_distance_f64) -> chi = smooth Heaviside(phi,
 * moll_zone) -> least-squares extrapolation of eta on the doubled grid
 * (sweeps until one fits nothing: lsq_sweeps queued launches, a longer
 * loop finished by one workgroup) -> solid stress
 * blended by chi -> omega += dt curl(div sigma) -> ENO3 -> RK2.  FP64.
 * The caller owns the state buffers ((nr, nz) doubles; scalars
 * AXS_RIGID_SCALARS doubles: t, dt, dt/dx, nu dt/2, nu dt, max|u|, step
 * counter, sweeps used in the last step, steps whose loop outran the
 * lsq_sweeps queued launches and was finished by the one-workgroup tail;
 * r (nr), z (nz), kill_wz (2 kill_width), kill_wr (kill_width)); the
 * stepper owns its scratch.  nz must equal 2 nr (the extrapolation uses z
 * as both axes of the doubled grid, as the reference does). */
typedef struct axs_soft_stepper axs_soft_stepper;
typedef struct {
  int nr, nz, kill_width;
  int lsq_sweeps;        /* extrapolation sweeps queued per step as launches */
  int reinit_iterations; /* < 0: 2 ceil(reinit_band / dx) + 8 */
  int lsq_grow;          /* nonzero: graph runs raise lsq_sweeps to the most sweeps a step
                            has needed so far + 4, rounded up to 4 (at most 4096), as they
                            learn it (read back without blocking; the graphs are captured
                            again); eager runs keep it */
  double dx, nu, G, rho_f, CFL, eps;
  double r_ball, z_cm, r_cm;               /* phi pin: r_ball - |eta - (z_cm, r_cm)| */
  double moll_zone, extrap_zone, reinit_band;
  /* streaming != 0: examples/SoftSphereStreaming/soft_sphere_streaming.py:128-264
   * instead of the Taylor-Green loop: omega is advected right after the level
   * set's reinit (before the solid stress acts); after the stress a rigid
   * core (radius tether_rad, centre (z_cm + e_amp r_ball sin(omega t), r_cm),
   * mask smoothed over moll_zone) is penalised (brink_lam) towards
   * U0 cos(omega t), u becoming the penalised velocity and omega gaining the
   * curl of the correction; dt is clipped to the end of the cycle
   * (freq_limit; scalars slot 9 holds the cycle timer, reset when it
   * reaches freq_limit) and to t_end.  The soft batch takes streaming = 0. */
  int streaming;
  double brink_lam, U0, omega, e_amp, tether_rad, freq_limit, t_end;
} axs_soft_config;
typedef struct {
  double *vorticity, *psi, *u_z, *u_r, *eta1, *eta2, *phi, *char_func, *scalars;
  const double *r, *z, *kill_wz, *kill_wr;
} axs_soft_buffers;
axs_status axs_soft_stepper_create(const axs_soft_config* cfg, const axs_soft_buffers* bufs,
                                   axs_soft_stepper** out);
axs_status axs_soft_stepper_destroy(axs_soft_stepper* st);
/* Queues nsteps steps on `stream` (use_graph: captured once, replayed). */
axs_status axs_soft_stepper_run(axs_soft_stepper* st, int nsteps, int use_graph, void* stream);

/* Soft batch: many soft-sphere runs (the loop of
 * examples/SoftSphereInTaylorGreenVortex/soft_sphere_in_taylor_green_vortex.py:101-213,
 * each with its own grid and solid) stepped together, every pass of the
 * step one launch over all of them.  The serving form of the soft step; the
 * reference runs one such loop per process.  A case's arithmetic is the
 * single-case step's with the any-length row transforms (FP64), so a case
 * gives the same bits alone and in any mix.  Cases join (add) and leave
 * (remove) between runs; both wait for the device to finish queued work.
 * The configuration and buffers are those of axs_soft_stepper_create
 * (lsq_sweeps and lsq_grow are the batch's: the queue starts at 12 sweeps
 * and grows as in lsq_grow); the buffers must stay valid until the case is
 * removed.  At most 1024 cases; nz up to 4096 (nr up to 2048). */
typedef struct axs_soft_batch axs_soft_batch;
axs_status axs_soft_batch_create(int max_cases, axs_soft_batch** out);
axs_status axs_soft_batch_add(axs_soft_batch* b, const axs_soft_config* cfg,
                              const axs_soft_buffers* bufs, int* case_id);
/* FP32-bulk soft batch (the serving form's precision; no single-process
 * counterpart in the reference, whose loop is FP64 throughout): the flow
 * fields vorticity, psi, u_z, u_r and char_func are floats, as are the
 * batch's spectrum, stress and RK2 scratch; the row transforms, velocity,
 * stress divergence, vorticity ENO3 and RK2 run in FP32.  FP64 stays for
 * the radial elimination (FP32 spectrum in, FP64 recurrences), the
 * reference map eta1/eta2 (its ENO3 step with the FP32 velocity
 * converted), the stress arithmetic (stored as float), the level set phi
 * and its reinit, the least-squares extrapolation and the scalars, so the
 * mask and the extrapolation's cells come from FP64 values.  Cases of one
 * batch share one storage type: an FP32 batch takes only
 * axs_soft_batch_add_f32 and an FP64 one only axs_soft_batch_add; remove,
 * run, info and destroy take either. */
typedef struct {
  float *vorticity, *psi, *u_z, *u_r;
  double *eta1, *eta2, *phi;
  float* char_func;
  double* scalars;
  const double *r, *z, *kill_wz, *kill_wr;
} axs_soft_buffers_f32;
axs_status axs_soft_batch_create_f32(int max_cases, axs_soft_batch** out);
axs_status axs_soft_batch_add_f32(axs_soft_batch* b, const axs_soft_config* cfg,
                                  const axs_soft_buffers_f32* bufs, int* case_id);
axs_status axs_soft_batch_remove(axs_soft_batch* b, int case_id);
/* Queues nsteps steps of every case on `stream` (use_graph: captured after
 * each change of the case set, replayed). */
axs_status axs_soft_batch_run(axs_soft_batch* b, int nsteps, int use_graph, void* stream);
/* Cases, extrapolation sweeps queued per step, graph launches and captures
 * so far (any pointer may be null). */
axs_status axs_soft_batch_info(const axs_soft_batch* b, int* ncase, int* queued_sweeps,
                               int64_t* graph_launches, int64_t* captures);
axs_status axs_soft_batch_destroy(axs_soft_batch* b);

/* ---- Device-resident step of examples/PeriodicRigidSlab/periodic_rigid_slab.py:93-158:
 * a rigid annular wall (char_func) oscillating along z at U0 cos(omega t),
 * Brinkman penalised, in a z-periodic domain of nz columns whose first and
 * last `ghost` columns are ghost copies; the periodic Stokes solve
 * (Neumann r) runs on the nz - 2 ghost interior columns.  Step order:
 * kill_boundary_vorticity_sine_r -> psi -> compute_velocity_from_psi_periodic
 * -> dt = min(0.9 dx^2 / 4 / nu, CFL dx / (max|u_z| + |u_r| + eps)) ->
 * brinkmann_penalize (u becomes the penalised velocity) ->
 * compute_vorticity_from_velocity_periodic of the correction, added to omega
 * -> advect_vorticity_via_eno3_periodic -> diffusion_RK2_periodic.  FP64.
 * Buffers: (nr, nz) vorticity, psi, u_z, u_r, char_func; scalars
 * AXS_RIGID_SCALARS doubles (t, dt, dt/dx, nu dt/2, nu dt, U0 cos(omega t),
 * max|u|, step counter); r (nr); kill_wr (kill_width weights of the top
 * rows, sin(pi (1 - r - dx/2) / 2 / (kill_width - 1) / dx) as the reference
 * forms them). */
typedef struct axs_slab_stepper axs_slab_stepper;
typedef struct {
  int nr, nz, ghost, kill_width;
  double dx, nu, lam, U0, omega, CFL, eps;
} axs_slab_config;
typedef struct {
  double *vorticity, *psi, *u_z, *u_r, *scalars;
  const double *char_func, *r, *kill_wr;
} axs_slab_buffers;
axs_status axs_slab_stepper_create(const axs_slab_config* cfg, const axs_slab_buffers* bufs,
                                   axs_slab_stepper** out);
axs_status axs_slab_stepper_destroy(axs_slab_stepper* st);
axs_status axs_slab_stepper_run(axs_slab_stepper* st, int nsteps, int use_graph, void* stream);

/* ---- Device-resident step of examples/PeriodicSoftSlab/periodic_soft_slab.py:214-396:
 * an elastic layer (solid_phi > 0, blended by solid_char_func) driven through
 * the fluid by an annular wall (wall_char_func) oscillating along z at
 * V_wall cos(omega t), z-periodic as the rigid slab (ghost columns, the
 * periodic Stokes solve on the interior columns).  Step order: dt = min(CFL
 * dx / sqrt(G / rho), 0.9 dx^2 / 4 / nu, CFL dx / (max|u_z| + |u_r| + eps)),
 * clipped to the sample timer (period freq_limit, starting at its limit, so
 * the first dt is 0) and to t_end -> kill_boundary_vorticity_sine_r -> psi
 * -> compute_velocity_from_psi_periodic (a sample, when the timer has run
 * out: t and u_z of rows < sample_rows in column nz / 2) ->
 * advect_refmap_via_eno3_periodic -> brinkmann_penalize towards the wall ->
 * compute_vorticity_from_velocity_periodic of the correction, added to omega
 * -> eta *= (solid_char_func > 0.5) -> the static PDE extrapolation of eta1
 * and eta2 (bounded_static_PDE_extrapolation.py:58-123, Jacobi to
 * extrap_tol, at most max_extrap_iterations per loop) ->
 * solid_sigma_periodic, blended -> update_vorticity_from_solid_stress_periodic
 * -> advect_vorticity_via_eno3_periodic -> diffusion_RK2_periodic.  FP64.
 * Buffers: (nr, nz) vorticity, psi, u_z, u_r, eta1, eta2, solid_phi (its
 * ghost columns are exchanged once), solid_char_func, wall_char_func; r (nr);
 * kill_wr as for the rigid slab; scalars AXS_RIGID_SCALARS doubles (t, dt,
 * dt/dx, nu dt/2, nu dt, V_wall cos(omega t), max|u|, step counter, sample
 * timer, samples taken, Jacobi iterations of the last step, loops that hit
 * the iteration bound); samples: sample_cap rows of 1 + sample_rows doubles
 * (t, then the profile). */
typedef struct axs_soft_slab_stepper axs_soft_slab_stepper;
typedef struct {
  int nr, nz, ghost, kill_width, sample_rows, sample_cap, max_extrap_iterations;
  double dx, nu, G, rho, lam, V_wall, omega, CFL, eps, z_max, extrap_tol, extrap_band,
      freq_limit, t_end;
} axs_soft_slab_config;
typedef struct {
  double *vorticity, *psi, *u_z, *u_r, *eta1, *eta2, *solid_phi, *scalars, *samples;
  const double *solid_char_func, *wall_char_func, *r, *kill_wr;
} axs_soft_slab_buffers;
axs_status axs_soft_slab_stepper_create(const axs_soft_slab_config* cfg,
                                        const axs_soft_slab_buffers* bufs,
                                        axs_soft_slab_stepper** out);
axs_status axs_soft_slab_stepper_destroy(axs_soft_slab_stepper* st);
axs_status axs_soft_slab_stepper_run(axs_soft_slab_stepper* st, int nsteps, int use_graph,
                                     void* stream);

/* ---- Device-resident step of examples/TorusStreaming/torus_streaming.py:129-233
 * with implicit diffusion (the driver's setting): a torus (tube radius a,
 * cross-section aspect AR, centre (Z_cm, R_cm(t)), R_cm(t) = R_cm + e a
 * sum_k amp_k sin(omega_k t)) dilating in r.  Step order: kill taper (z
 * then r) -> psi -> u from psi -> the torus mask at R_cm(t) -> potential of
 * char_func * rate / r, rate = e a sum_k amp_k omega_k cos(omega_k t), u +=
 * grad phi -> brinkmann_penalize towards (0, rate) into u_z_pen, u_r_pen ->
 * omega += curl(u_pen - u) -> ENO3 with u -> implicit Euler diffusion at the
 * fixed dt.  FP64.  Buffers: (nr, nz) vorticity, psi, u_z, u_r, u_z_pen,
 * u_r_pen, char_func, vel_phi; scalars AXS_RIGID_SCALARS doubles (t, dt,
 * dt/dx, R_cm(t), rate, step counter); r (nr), z (nz), kill_wz
 * (2 kill_width), kill_wr (kill_width) as for the rigid stepper.  The
 * potential's singular mode is solved in the gauge of
 * axs_potential_solve_f64 (the velocity does not see it). */
typedef struct axs_torus_stepper axs_torus_stepper;
typedef struct {
  int nr, nz, kill_width, nmodes; /* nmodes <= 4 */
  double dx, nu, lam, moll_zone, a, AR, e, Z_cm, R_cm, dt;
  double amp[4], omega[4];
} axs_torus_config;
typedef struct {
  double *vorticity, *psi, *u_z, *u_r, *u_z_pen, *u_r_pen, *char_func, *vel_phi, *scalars;
  const double *r, *z, *kill_wz, *kill_wr;
} axs_torus_buffers;
axs_status axs_torus_stepper_create(const axs_torus_config* cfg, const axs_torus_buffers* bufs,
                                    axs_torus_stepper** out);
axs_status axs_torus_stepper_destroy(axs_torus_stepper* st);
axs_status axs_torus_stepper_run(axs_torus_stepper* st, int nsteps, int use_graph, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* AXISYM_H_ */
