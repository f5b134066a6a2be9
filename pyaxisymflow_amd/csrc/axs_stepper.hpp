// Rigid-body stepper state (shared by the C ABI and its kernels).
#pragma once

#include "axs_stokes.hpp"

namespace axs {

// Device scalar slots (double array of kScCount entries).
enum : int {
  kScT = 0,       // time
  kScDt,          // dt of the current step
  kScDtByDx,      // dt / dx
  kScLamDt,       // brink_lam * dt
  kScHalfNuDt,    // 0.5 * nu * dt
  kScNuDt,        // nu * dt
  kScPrefac,      // free-stream ramp factor for the current step
  kScUpen,        // penalisation velocity U_z(t) for the current step
  kScMaxU,        // max(|u_z| + |u_r|)
  kScDragSum,     // sum(r * chi * u_pen_z)
  kScIt,          // step counter
  kScCount
};

struct RigidConfig {
  int nr, nz, kill_width, history;
  double dx, nu, brink_lam, U0, T_ramp, CFL, eps, dt_diffusive;
  double osc_amp, osc_omega, drag_scale, drag_div;
};

struct RigidBuffers {
  double *vorticity, *psi, *u_z, *u_r, *u_pen_z, *u_pen_r, *char_func, *work0, *work1;
  double* scalars;
  double* drag;
  const double* r;
  const double* kill_wz;
  const double* kill_wr;
};

constexpr int kRigidStages = 10;
extern const char* const kRigidStageNames[kRigidStages];

struct RigidStepper {
  RigidConfig cfg{};
  RigidBuffers buf{};
  StokesPlan* plan = nullptr;
  double* partials = nullptr;
  int blocks_x = 0, npartials = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  hipStream_t graph_stream = nullptr;

  hipError_t init(const RigidConfig& c, const RigidBuffers& b);
  hipError_t enqueue_stage(int stage, hipStream_t s);
  hipError_t enqueue_step(hipStream_t s);
  hipError_t run(int nsteps, bool use_graph, hipStream_t s);
  ~RigidStepper();
};

}  // namespace axs
