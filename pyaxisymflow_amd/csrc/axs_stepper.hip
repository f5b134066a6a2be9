// Device-resident timestep of the rigid-body drivers (row A16):
// examples/FlowPastSphere/flow_past_sphere.py:107-188 (static body, free
// stream with sine ramp, drag coefficient) and
// examples/SpheroidStreaming/spheroid_streaming.py:129-246 (oscillating
// spheroid in quiescent fluid: mask rebuilt every step, penalisation
// velocity U_0 cos(omega t), dt clipped to the period and to tEnd, psi
// averaged over each period).
//
// One step, in the reference's order, as seven passes (a stage is skipped
// when its feature is off):
//   taper_dct2_rows   kill taper of the stored field on the fly, DCT-II of
//                     r * omega along z
//   tridiag_columns   one radial tridiagonal solve per z-mode
//   dct3_rows         inverse DCT -> psi
//   velocity_max      u from psi (+ free stream), max|u| (atomicMax slots)
//   body_mask         chi = H(phi(t)) of the moving body
//   step_fused        dt rule, u from psi, Brinkman, taper + curl(u_pen - u),
//                     force partials, psi average, ENO3, both RK2 stages ->
//                     the other buffer of the vorticity pair
//   finalize_step     force record, t, period timer, step count
// The stored vorticity alternates between the home buffer and work0; one
// graph is captured per direction, and a run that ends on work0 copies it
// home, so the home buffer holds the field whenever a run returns.
// Every scalar (t, dt, ramp factor, penalisation velocity, period timer,
// force history) stays in device memory, so a step is captured once into a
// hipGraph and replayed.  No memset nodes are recorded (fills are kernels).
//
// Precision: T = double is the plain FP64 path.  T = float stores every
// field in FP32 and runs the stencils and z-transforms in FP32, while the
// radial tridiagonal elimination, the mask level set, and all reductions
// and step scalars stay FP64.
#include "axs_fused.hpp"
#include "axs_stepper.hpp"

#include <new>

namespace axs {

// ------------------------------------------------------------------ kernels

// chi = smooth_Heaviside(phi) of the spheroid centred at z = sc[BodyZ]
// (spheroid_streaming.py:207-212), level set and blend in FP64.
template <typename T>
__global__ void body_mask_kernel(T* __restrict__ chi, const double* __restrict__ r,
                                 const double* __restrict__ z, int nr, int nz,
                                 const double* __restrict__ sc, RigidConfig cfg) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  const double zc = sc[kScBodyZ];
  const double dz = z[j] - zc, dr = r[i] - cfg.body_r0;
  const double phi = cfg.body_radius - sqrt(dz * dz / cfg.body_aspect2 + dr * dr);
  const double w = cfg.moll_zone;
  double h = (phi >= w) ? 1.0 : 0.0;
  if (fabs(phi) < w) h = h + 0.5 * (1.0 + phi / w + sin(M_PI * phi / w) / M_PI);
  chi[(int64_t)i * nz + j] = T(h);
}

// Step scalars for the time now in sc[T] (without advancing); clears the
// velocity maximum slots.
__global__ void init_scalars_kernel(double* __restrict__ sc, double* __restrict__ slots,
                                    RigidConfig cfg) {
  if (threadIdx.x < kMaxSlots) slots[threadIdx.x * kSlotStride] = 0.0;
  if (threadIdx.x == 0) next_step_scalars(sc, cfg, sc[kScT]);
}

// ------------------------------------------------------------------ stepper

const char* const kRigidStageNames[kRigidStages] = {
    "taper_dct2_rows", "tridiag_columns", "dct3_rows",    "velocity_max",
    "body_mask",       "step_fused",      "finalize_step"};

template <typename T>
__global__ void copy_field_kernel(T* __restrict__ dst, const T* __restrict__ src, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[k];
}

RigidStepperBase::~RigidStepperBase() {
  for (int q = 0; q < 2; ++q) {
    if (graph_exec[q]) (void)hipGraphExecDestroy(graph_exec[q]);
    if (graph[q]) (void)hipGraphDestroy(graph[q]);
  }
  if (partials) (void)hipFree(partials);
  if (slots) (void)hipFree(slots);
  delete plan;
}

hipError_t RigidStepperBase::init(const RigidConfig& c, const RigidBuffers& b) {
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
  npartials = fused_blocks(cfg.nr, cfg.nz);
  e = hipMalloc(reinterpret_cast<void**>(&partials), sizeof(double) * (size_t)npartials);
  if (e != hipSuccess) return e;
  e = hipMalloc(reinterpret_cast<void**>(&slots), sizeof(double) * kMaxSlots * kSlotStride);
  if (e != hipSuccess) return e;
  init_scalars_kernel<<<1, 64>>>(buf.scalars, slots, cfg);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipDeviceSynchronize();
}

bool RigidStepperBase::stage_enabled(int stage) const {
  if (stage == 4) return cfg.moving_body != 0;
  return stage >= 0 && stage < kRigidStages;
}

template <typename T>
hipError_t RigidStepper<T>::enqueue_stage(int stage, hipStream_t s) {
  const int nr = cfg.nr, nz = cfg.nz, w = cfg.kill_width;
  double* sc = buf.scalars;
  T* home = static_cast<T*>(buf.vorticity);
  T* other = static_cast<T*>(buf.work0);
  T* cur = parity ? other : home;
  T* nxt = parity ? home : other;
  T* psi = static_cast<T*>(buf.psi);
  T* chi = static_cast<T*>(buf.char_func);
  T* spec = reinterpret_cast<T*>(plan->spec);
  if (!stage_enabled(stage)) return hipSuccess;
  switch (stage) {
    case 0:
      return launch_taper_dct2_rows<T>(*plan, cur, nullptr, spec, buf.kill_wz, buf.kill_wr, w, s);
    case 1:
      return launch_tridiag<T>(*plan, spec, s);
    case 2:
      return launch_dct3_rows<T>(*plan, spec, psi, nr, s);
    case 3:
      return launch_velocity_max<T>(psi, buf.r, nr, nz, T(cfg.dx), sc, slots, cfg, s);
    case 4:
      body_mask_kernel<T><<<dim3(blocks_x, nr), 256, 0, s>>>(chi, buf.r, buf.z, nr, nz, sc, cfg);
      return hipGetLastError();
    case 5:
      return launch_step_fused<T>(psi, chi, cur, nxt,
                                  cfg.average_psi ? static_cast<T*>(buf.avg_psi) : nullptr,
                                  buf.r, buf.kill_wz, buf.kill_wr, w, nr, nz, T(cfg.dx), sc, slots,
                                  partials, cfg, s);
    case 6:
      return launch_finalize_step(partials, npartials, slots, sc, buf.drag, cfg, s);
    default:
      return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t RigidStepper<T>::settle(hipStream_t s) {
  if (!parity) return hipSuccess;
  const int64_t n = (int64_t)cfg.nr * cfg.nz;
  copy_field_kernel<T><<<ceil_div(n, 256), 256, 0, s>>>(static_cast<T*>(buf.vorticity),
                                                        static_cast<const T*>(buf.work0), n);
  parity = 0;
  return hipGetLastError();
}

template <typename T>
hipError_t RigidStepper<T>::velocity_out(hipStream_t s) {
  return launch_velocity_out<T>(static_cast<const T*>(buf.psi),
                                static_cast<const T*>(buf.char_func),
                                static_cast<T*>(buf.u_pen_z), static_cast<T*>(buf.u_pen_r), buf.r,
                                cfg.nr, cfg.nz, T(cfg.dx), buf.scalars, cfg, s);
}

hipError_t RigidStepperBase::enqueue_step(hipStream_t s) {
  for (int q = 0; q < kRigidStages; ++q) {
    hipError_t e = enqueue_stage(q, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t RigidStepperBase::run(int nsteps, bool use_graph, hipStream_t s) {
  if (!use_graph) {
    for (int q = 0; q < nsteps; ++q) {
      hipError_t e = enqueue_step(s);
      if (e != hipSuccess) return e;
      parity ^= 1;
    }
    return settle(s);
  }
  if (graph_stream != s) {
    for (int q = 0; q < 2; ++q) {
      if (graph_exec[q]) (void)hipGraphExecDestroy(graph_exec[q]);
      if (graph[q]) (void)hipGraphDestroy(graph[q]);
      graph_exec[q] = nullptr;
      graph[q] = nullptr;
    }
    const int saved = parity;
    for (int q = 0; q < 2; ++q) {
      parity = q;
      hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) {
        parity = saved;
        return e;
      }
      hipError_t le = enqueue_step(s);
      e = hipStreamEndCapture(s, &graph[q]);
      if (le == hipSuccess && e == hipSuccess)
        e = hipGraphInstantiate(&graph_exec[q], graph[q], nullptr, nullptr, 0);
      if (le != hipSuccess || e != hipSuccess) {
        parity = saved;
        return le != hipSuccess ? le : e;
      }
    }
    parity = saved;
    graph_stream = s;
  }
  for (int q = 0; q < nsteps; ++q) {
    hipError_t e = hipGraphLaunch(graph_exec[parity], s);
    if (e != hipSuccess) return e;
    parity ^= 1;
  }
  return settle(s);
}

template struct RigidStepper<double>;
template struct RigidStepper<float>;

}  // namespace axs

using namespace axs;

struct axs_rigid_stepper {
  RigidStepperBase* impl = nullptr;
  ~axs_rigid_stepper() { delete impl; }
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
  AXS_CHECK_ARG(cfg->precision == AXS_F64 || cfg->precision == AXS_F32, "bad precision");
  AXS_CHECK_ARG(!cfg->moving_body || (cfg->body_aspect2 > 0 && cfg->moll_zone > 0),
                "moving body needs aspect^2 > 0 and moll_zone > 0");
  AXS_CHECK_ARG(bufs->vorticity && bufs->psi && bufs->u_z && bufs->u_r && bufs->u_pen_z &&
                    bufs->u_pen_r && bufs->char_func && bufs->work0 && bufs->work1 &&
                    bufs->scalars && bufs->r && bufs->kill_wz && bufs->kill_wr,
                "null buffer");
  AXS_CHECK_ARG(!cfg->average_psi || bufs->avg_psi, "average_psi needs the avg_psi buffer");
  AXS_CHECK_ARG(!cfg->moving_body || bufs->z, "moving body needs the z table");
  axs_rigid_stepper* st = new (std::nothrow) axs_rigid_stepper();
  if (!st) {
    set_error("out of host memory");
    return AXS_ERR_ALLOC;
  }
  if (cfg->precision == AXS_F64)
    st->impl = new (std::nothrow) RigidStepper<double>();
  else
    st->impl = new (std::nothrow) RigidStepper<float>();
  if (!st->impl) {
    delete st;
    set_error("out of host memory");
    return AXS_ERR_ALLOC;
  }
  RigidConfig c{};
  c.nr = cfg->nr;
  c.nz = cfg->nz;
  c.kill_width = cfg->kill_width;
  c.history = cfg->history;
  c.precision = cfg->precision;
  c.moving_body = cfg->moving_body;
  c.average_psi = cfg->average_psi;
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
  c.dt_cap = cfg->dt_cap;
  c.period = cfg->period;
  c.t_end = cfg->t_end;
  c.body_z0 = cfg->body_z0;
  c.body_r0 = cfg->body_r0;
  c.body_aspect2 = cfg->body_aspect2;
  c.body_radius = cfg->body_radius;
  c.body_disp = cfg->body_disp;
  c.moll_zone = cfg->moll_zone;
  RigidBuffers b{bufs->vorticity, bufs->psi,    bufs->u_z,     bufs->u_r,    bufs->u_pen_z,
                 bufs->u_pen_r,   bufs->char_func, bufs->work0, bufs->work1,  bufs->avg_psi,
                 bufs->scalars,   bufs->drag,   bufs->r,       bufs->z,      bufs->kill_wz,
                 bufs->kill_wr};
  hipError_t e = st->impl->init(c, b);
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
  AXS_CHECK_ARG(!use_graph || stream, "graph capture needs a non-null stream");
  AXS_HIP(st->impl->run(nsteps, use_graph != 0, as_stream(stream)));
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
  for (int q = 0; q < reps; ++q) AXS_HIP(st->impl->enqueue_stage(stage, as_stream(stream)));
  return AXS_OK;
}

axs_status axs_rigid_stepper_velocity(axs_rigid_stepper* st, void* stream) {
  AXS_CHECK_ARG(st, "null stepper");
  AXS_HIP(st->impl->velocity_out(as_stream(stream)));
  return AXS_OK;
}

axs_status axs_rigid_stepper_reset_scalars(axs_rigid_stepper* st, void* stream) {
  AXS_CHECK_ARG(st, "null stepper");
  init_scalars_kernel<<<1, 64, 0, as_stream(stream)>>>(st->impl->buf.scalars, st->impl->slots,
                                                       st->impl->cfg);
  AXS_LAUNCH_CHECK("init_scalars");
  return AXS_OK;
}

}  // extern "C"
