// Device-resident timestep of the rigid-body drivers (row A16):
// examples/FlowPastSphere/flow_past_sphere.py:107-188 (static body, free
// stream with sine ramp, drag coefficient) and, with an oscillating
// penalisation velocity, examples/SpheroidStreaming/spheroid_streaming.py:129-246.
//
// One step, in the reference's order:
//   kill taper (z, then r)                       kill_z_tab / kill_r_tab
//   psi = Stokes solve(omega)                    dct2 rows, tridiagonal columns, dct3 rows
//   u = velocity(psi) (+ free stream), max|u|    velocity_freestream
//   dt = min(diffusive, CFL dx / (max + eps))    finalize_dt (one block)
//   u <- Brinkman(u); omega += curl(u - u_raw);
//   drag partial sums                            penalize_curl
//   Cd, t += dt, it += 1                         finalize_step (one block)
//   omega <- ENO3(omega, u)                      eno3 (out of place)
//   omega <- RK2 diffusion                       rk2 stage 1, stage 2 (back home)
// Every scalar (t, dt, ramp factor, penalisation velocity, drag history)
// stays in device memory, so a step can be captured once into a hipGraph and
// replayed.  No memset nodes are recorded (fills are kernels).
#include "axs_eno3.hpp"
#include "axs_reduce.hpp"
#include "axs_stokes.hpp"
#include "axs_stepper.hpp"

#include <new>

namespace axs {

template <typename T, class RA>
hipError_t launch_rk2(const T* w_in, T* tmp, T* w_out, RA R, int nr, int nz, Scalar<T> half_nu_dt,
                      Scalar<T> nu_dt, T dx, hipStream_t s);

// ------------------------------------------------------------------ kernels

// kill_boundary_vorticity_sine_z with tabulated weights (wz[0..w) left,
// wz[w..2w) right; the host forms them with the reference's expression).
template <typename T>
__global__ void kill_z_tab_kernel(T* __restrict__ w, const double* __restrict__ wz, int nr,
                                  int nz, int width) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  T* row = w + (int64_t)i * nz;
  const T left = row[width - 1];
  for (int j = 0; j < width; ++j) row[j] = T(wz[j]) * left;
  const T right = row[nz - width];
  for (int j = 0; j < width; ++j) row[nz - width + j] = T(wz[width + j]) * right;
}

// kill_boundary_vorticity_sine_r with tabulated weights wr[0..w) for rows
// nr-w..nr-1; then the axis row is zeroed.
template <typename T>
__global__ void kill_r_tab_kernel(T* __restrict__ w, const double* __restrict__ wr, int nr,
                                  int nz, int width) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nz) return;
  const T src = w[(int64_t)(nr - width) * nz + j];
  for (int q = 0; q < width; ++q) w[(int64_t)(nr - width + q) * nz + j] = T(wr[q]) * src;
  w[j] = T(0);
}

template <typename T>
__device__ __forceinline__ double block_max(double v) {
  __shared__ double sm[256];
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + off]);
    __syncthreads();
  }
  return sm[0];
}

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sm[256];
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) sm[threadIdx.x] = sm[threadIdx.x] + sm[threadIdx.x + off];
    __syncthreads();
  }
  return sm[0];
}

// compute_velocity_from_psi_unb + u_z += U0 * prefac + per-block max(|u_z|+|u_r|).
// Block = 256 columns of one row; partial index = row * gridDim.x + blockIdx.x.
template <typename T>
__global__ void __launch_bounds__(256) velocity_freestream_kernel(
    T* __restrict__ u_z, T* __restrict__ u_r, const T* __restrict__ psi,
    const double* __restrict__ r, int nr, int nz, T dx, const double* __restrict__ sc,
    double U0, double* __restrict__ partials) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  double m = -1.0;
  if (j < nz) {
    const int64_t c = (int64_t)i * nz + j;
    const T two_dx = T(2) * dx;
    const T ri = T(r[i]);
    T uz;
    if (i == 0)
      uz = (-psi[c + 2 * nz] + T(4) * psi[c + nz] - T(3) * psi[c]) / two_dx / ri;
    else if (i == nr - 1)
      uz = (psi[c - 2 * nz] - T(4) * psi[c - nz] + T(3) * psi[c]) / two_dx / ri;
    else
      uz = (psi[c + nz] - psi[c - nz]) / two_dx / ri;
    T ur;
    if (j == 0)
      ur = -(-psi[c + 2] + T(4) * psi[c + 1] - T(3) * psi[c]) / two_dx / ri;
    else if (j == nz - 1)
      ur = -(psi[c - 2] - T(4) * psi[c - 1] + T(3) * psi[c]) / two_dx / ri;
    else
      ur = -(psi[c + 1] - psi[c - 1]) / two_dx / ri;
    uz = uz + T(U0 * sc[kScPrefac]);
    u_z[c] = uz;
    u_r[c] = ur;
    m = double(fabs(uz) + fabs(ur));
  }
  const double b = block_max<T>(m);
  if (threadIdx.x == 0) partials[(int64_t)i * gridDim.x + blockIdx.x] = b;
}

// dt rule (flow_past_sphere.py:150-153) and the derived step scalars.
__global__ void __launch_bounds__(256) finalize_dt_kernel(const double* __restrict__ partials,
                                                          int count, double* __restrict__ sc,
                                                          RigidConfig cfg) {
  double m = -1.0;
  for (int q = threadIdx.x; q < count; q += 256) m = fmax(m, partials[q]);
  m = block_max<double>(m);
  if (threadIdx.x == 0) {
    const double cfl_dt = cfg.CFL * cfg.dx / (m + cfg.eps);
    const double dt = (cfl_dt < cfg.dt_diffusive) ? cfl_dt : cfg.dt_diffusive;
    sc[kScMaxU] = m;
    sc[kScDt] = dt;
    sc[kScDtByDx] = dt / cfg.dx;
    sc[kScLamDt] = cfg.brink_lam * dt;
    sc[kScHalfNuDt] = 0.5 * cfg.nu * dt;
    sc[kScNuDt] = cfg.nu * dt;
  }
}

// brinkmann_penalize (U_z = sc[U_pen], U_r = 0) into (pz, pr); omega's
// interior += curl(u_pen - u_raw) (compute_vorticity_from_velocity_unb on
// the difference); per-block partial of sum(r * chi * u_pen_z).
template <typename T>
__global__ void __launch_bounds__(256) penalize_curl_kernel(
    const T* __restrict__ uz, const T* __restrict__ ur, const T* __restrict__ chi,
    T* __restrict__ pz, T* __restrict__ pr, T* __restrict__ w, const double* __restrict__ r,
    int nr, int nz, T dx, const double* __restrict__ sc, double* __restrict__ partials) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  double acc = 0.0;
  if (j < nz) {
    const T lam_dt = T(sc[kScLamDt]);
    const T U = T(sc[kScUpen]);
    const int64_t c = (int64_t)i * nz + j;
    auto pen = [&](int64_t q, T u, T Uv) {
      const T ldc = lam_dt * chi[q];
      return (u + ldc * Uv) / (T(1) + ldc);
    };
    const T uzc = uz[c], urc = ur[c];
    const T pzc = pen(c, uzc, U);
    pz[c] = pzc;
    pr[c] = pen(c, urc, T(0));
    if (i >= 1 && i <= nr - 2 && j >= 1 && j <= nz - 2) {
      const T two_dx = T(2) * dx;
      const T dr_p = pen(c + 1, ur[c + 1], T(0)) - ur[c + 1];
      const T dr_m = pen(c - 1, ur[c - 1], T(0)) - ur[c - 1];
      const T dz_p = pen(c + nz, uz[c + nz], U) - uz[c + nz];
      const T dz_m = pen(c - nz, uz[c - nz], U) - uz[c - nz];
      const T pv = (dr_p - dr_m) / two_dx - (dz_p - dz_m) / two_dx;
      w[c] = w[c] + pv;
    }
    acc = (r[i] * double(chi[c])) * double(pzc);
  }
  const double b = block_sum(acc);
  if (threadIdx.x == 0) partials[(int64_t)i * gridDim.x + blockIdx.x] = b;
}

// Drag coefficient into the history, then t += dt, it += 1, and the next
// step's ramp factor and penalisation velocity.
__global__ void __launch_bounds__(256) finalize_step_kernel(const double* __restrict__ partials,
                                                            int count, double* __restrict__ sc,
                                                            double* __restrict__ drag,
                                                            RigidConfig cfg) {
  double s = 0.0;
  for (int q = threadIdx.x; q < count; q += 256) s += partials[q];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    const long it = (long)sc[kScIt];
    if (drag && it < cfg.history) drag[it] = (cfg.drag_scale * s) / cfg.drag_div;
    sc[kScDragSum] = s;
    const double t = sc[kScT] + sc[kScDt];
    sc[kScT] = t;
    sc[kScIt] = double(it + 1);
    sc[kScPrefac] = (cfg.T_ramp > 0.0 && t < cfg.T_ramp) ? sin(0.5 * M_PI * t / cfg.T_ramp) : 1.0;
    sc[kScUpen] = cfg.osc_amp * cos(cfg.osc_omega * t);
  }
}

// Initial scalars for t = sc[T] (prefac, U_pen) without advancing.
__global__ void init_scalars_kernel(double* __restrict__ sc, RigidConfig cfg) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const double t = sc[kScT];
    sc[kScPrefac] = (cfg.T_ramp > 0.0 && t < cfg.T_ramp) ? sin(0.5 * M_PI * t / cfg.T_ramp) : 1.0;
    sc[kScUpen] = cfg.osc_amp * cos(cfg.osc_omega * t);
  }
}

// ------------------------------------------------------------------ stepper

RigidStepper::~RigidStepper() {
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (partials) (void)hipFree(partials);
  delete plan;
}

hipError_t RigidStepper::init(const RigidConfig& c, const RigidBuffers& b) {
  cfg = c;
  buf = b;
  plan = new (std::nothrow) StokesPlan();
  if (!plan) return hipErrorOutOfMemory;
  FftPlanArgs fp{};
  fp.M = cfg.nz / 2;
  fp.nstages = 0;
  int m = fp.M;
  for (int rdx : {4, 2, 3, 5})
    while (m % rdx == 0 && fp.nstages < kMaxFftStages) {
      fp.radix[fp.nstages++] = rdx;
      m /= rdx;
    }
  if (m != 1) return hipErrorInvalidValue;
  plan->fft = fp;
  hipError_t e = plan->init(cfg.nr, cfg.nz, cfg.dx);
  if (e != hipSuccess) return e;
  blocks_x = ceil_div(cfg.nz, 256);
  npartials = blocks_x * cfg.nr;
  e = hipMalloc(reinterpret_cast<void**>(&partials), sizeof(double) * (size_t)npartials * 2);
  if (e != hipSuccess) return e;
  init_scalars_kernel<<<1, 64>>>(buf.scalars, cfg);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipDeviceSynchronize();
}

// Stage table of one step (also used to time each stage on its own).
const char* const kRigidStageNames[kRigidStages] = {
    "kill_taper", "dct2_rows", "tridiag_columns", "dct3_rows", "velocity_freestream",
    "finalize_dt", "penalize_curl", "finalize_step", "eno3", "rk2_diffusion"};

hipError_t RigidStepper::enqueue_stage(int stage, hipStream_t s) {
  const int nr = cfg.nr, nz = cfg.nz, w = cfg.kill_width;
  double* sc = buf.scalars;
  double* part_u = partials;
  double* part_d = partials + npartials;
  const dim3 g2(blocks_x, nr);
  switch (stage) {
    case 0:
      kill_z_tab_kernel<double><<<ceil_div(nr, 64), 64, 0, s>>>(buf.vorticity, buf.kill_wz, nr, nz, w);
      kill_r_tab_kernel<double><<<ceil_div(nz, 256), 256, 0, s>>>(buf.vorticity, buf.kill_wr, nr, nz, w);
      return hipGetLastError();
    case 1:
      return launch_dct2_rows<double>(*plan, buf.vorticity, plan->spec, plan->rvec, nr, s);
    case 2:
      return launch_tridiag<double>(*plan, plan->spec, s);
    case 3:
      return launch_dct3_rows<double>(*plan, plan->spec, buf.psi, nr, s);
    case 4:
      velocity_freestream_kernel<double><<<g2, 256, 0, s>>>(buf.u_z, buf.u_r, buf.psi, buf.r, nr,
                                                            nz, cfg.dx, sc, cfg.U0, part_u);
      return hipGetLastError();
    case 5:
      finalize_dt_kernel<<<1, 256, 0, s>>>(part_u, npartials, sc, cfg);
      return hipGetLastError();
    case 6:
      penalize_curl_kernel<double><<<g2, 256, 0, s>>>(buf.u_z, buf.u_r, buf.char_func,
                                                      buf.u_pen_z, buf.u_pen_r, buf.vorticity,
                                                      buf.r, nr, nz, cfg.dx, sc, part_d);
      return hipGetLastError();
    case 7:
      finalize_step_kernel<<<1, 256, 0, s>>>(part_d, npartials, sc, buf.drag, cfg);
      return hipGetLastError();
    case 8:
      return launch_eno3_conservative<double>(buf.vorticity, buf.u_pen_z, buf.u_pen_r, buf.work0,
                                              nr, nz, Scalar<double>{0.0, sc + kScDtByDx}, -1.0, s);
    case 9:
      return launch_rk2<double, RRow<double>>(buf.work0, buf.work1, buf.vorticity,
                                              RRow<double>{buf.r}, nr, nz,
                                              Scalar<double>{0.0, sc + kScHalfNuDt},
                                              Scalar<double>{0.0, sc + kScNuDt}, cfg.dx, s);
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t RigidStepper::enqueue_step(hipStream_t s) {
  for (int q = 0; q < kRigidStages; ++q) {
    hipError_t e = enqueue_stage(q, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t RigidStepper::run(int nsteps, bool use_graph, hipStream_t s) {
  if (!use_graph) {
    for (int q = 0; q < nsteps; ++q) {
      hipError_t e = enqueue_step(s);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  if (!graph_exec || graph_stream != s) {
    if (graph_exec) {
      (void)hipGraphExecDestroy(graph_exec);
      graph_exec = nullptr;
    }
    if (graph) {
      (void)hipGraphDestroy(graph);
      graph = nullptr;
    }
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t le = enqueue_step(s);
    e = hipStreamEndCapture(s, &graph);
    if (le != hipSuccess) return le;
    if (e != hipSuccess) return e;
    e = hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return e;
    graph_stream = s;
  }
  for (int q = 0; q < nsteps; ++q) {
    hipError_t e = hipGraphLaunch(graph_exec, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace axs

using namespace axs;

struct axs_rigid_stepper {
  RigidStepper impl;
};

extern "C" {

axs_status axs_rigid_stepper_create(const axs_rigid_config* cfg, const axs_rigid_buffers* bufs,
                                    axs_rigid_stepper** out) {
  AXS_CHECK_ARG(cfg && bufs && out, "null pointer");
  *out = nullptr;
  AXS_CHECK_ARG(cfg->nr >= 4 && cfg->nz >= 8 && cfg->nz % 2 == 0, "need nr >= 4 and even nz >= 8");
  AXS_CHECK_ARG(cfg->kill_width >= 2 && 2 * cfg->kill_width <= cfg->nz &&
                    cfg->kill_width < cfg->nr,
                "bad kill_width for the grid");
  AXS_CHECK_ARG(cfg->dx > 0 && cfg->nu > 0 && cfg->CFL > 0, "dx, nu and CFL must be positive");
  AXS_CHECK_ARG(bufs->vorticity && bufs->psi && bufs->u_z && bufs->u_r && bufs->u_pen_z &&
                    bufs->u_pen_r && bufs->char_func && bufs->work0 && bufs->work1 &&
                    bufs->scalars && bufs->r && bufs->kill_wz && bufs->kill_wr,
                "null buffer");
  axs_rigid_stepper* st = new (std::nothrow) axs_rigid_stepper();
  if (!st) {
    set_error("out of host memory");
    return AXS_ERR_ALLOC;
  }
  RigidConfig c;
  c.nr = cfg->nr;
  c.nz = cfg->nz;
  c.kill_width = cfg->kill_width;
  c.history = cfg->history;
  c.dx = cfg->dx;
  c.nu = cfg->nu;
  c.brink_lam = cfg->brink_lam;
  c.U0 = cfg->U0;
  c.T_ramp = cfg->T_ramp;
  c.CFL = cfg->CFL;
  c.eps = cfg->eps;
  c.dt_diffusive = cfg->dt_diffusive;
  c.osc_amp = cfg->osc_amp;
  c.osc_omega = cfg->osc_omega;
  c.drag_scale = cfg->drag_scale;
  c.drag_div = cfg->drag_div;
  RigidBuffers b{bufs->vorticity, bufs->psi,     bufs->u_z,    bufs->u_r,
                 bufs->u_pen_z,   bufs->u_pen_r, bufs->char_func, bufs->work0,
                 bufs->work1,     bufs->scalars, bufs->drag,   bufs->r,
                 bufs->kill_wz,   bufs->kill_wr};
  hipError_t e = st->impl.init(c, b);
  if (e != hipSuccess) {
    set_error(std::string("rigid stepper: ") + hipGetErrorString(e));
    delete st;
    return AXS_ERR_HIP;
  }
  *out = st;
  return AXS_OK;
}

axs_status axs_rigid_stepper_destroy(axs_rigid_stepper* st) {
  delete st;
  return AXS_OK;
}

axs_status axs_rigid_stepper_run(axs_rigid_stepper* st, int nsteps, int use_graph, void* stream) {
  AXS_CHECK_ARG(st, "null stepper");
  AXS_CHECK_ARG(nsteps >= 0, "negative step count");
  AXS_HIP(st->impl.run(nsteps, use_graph != 0, as_stream(stream)));
  return AXS_OK;
}

int axs_rigid_stepper_stage_count(void) { return kRigidStages; }

const char* axs_rigid_stepper_stage_name(int stage) {
  return (stage >= 0 && stage < kRigidStages) ? kRigidStageNames[stage] : "";
}

axs_status axs_rigid_stepper_run_stage(axs_rigid_stepper* st, int stage, int reps, void* stream) {
  AXS_CHECK_ARG(st, "null stepper");
  AXS_CHECK_ARG(stage >= 0 && stage < kRigidStages, "bad stage");
  AXS_CHECK_ARG(reps >= 0, "negative repetition count");
  for (int q = 0; q < reps; ++q) AXS_HIP(st->impl.enqueue_stage(stage, as_stream(stream)));
  return AXS_OK;
}

axs_status axs_rigid_stepper_reset_scalars(axs_rigid_stepper* st, void* stream) {
  AXS_CHECK_ARG(st, "null stepper");
  init_scalars_kernel<<<1, 64, 0, as_stream(stream)>>>(st->impl.buf.scalars, st->impl.cfg);
  AXS_LAUNCH_CHECK("init_scalars");
  return AXS_OK;
}

}  // extern "C"
