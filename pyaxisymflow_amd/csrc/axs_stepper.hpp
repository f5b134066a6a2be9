// Rigid-body stepper state (shared by the C ABI and its kernels).
#pragma once

#include "axs_stokes.hpp"

#include <vector>

namespace axs {

// Device scalar slots (double array of kScCount entries).
enum : int {
  kScT = 0,       // time at the start of the current step
  kScDt,          // dt of the current step
  kScDtByDx,      // dt / dx
  kScLamDt,       // brink_lam * dt
  kScHalfNuDt,    // 0.5 * nu * dt
  kScNuDt,        // nu * dt
  kScPrefac,      // free-stream ramp factor for the current step
  kScUpen,        // penalisation velocity U_z(t) for the current step
  kScMaxU,        // max(|u_z| + |u_r|)
  kScDragSum,     // sum(r * chi * (u_pen_z - U_pen))
  kScIt,          // step counter
  kScFreqTimer,   // time since the last period boundary
  kScAvgReset,    // 1 when the psi average restarts this step
  kScBodyZ,       // centre of the moving body for the current step
  kScFreqStep,    // period timer after this step's reset test (before + dt)
  kScPrefacStep,  // ramp factor the last completed step used
  kScUpenStep,    // penalisation velocity the last completed step used
  kScPending,     // 1 while a step's force partials await finalisation
  kScCount
};

static_assert(kScCount <= AXS_RIGID_SCALARS, "scalar slots exceed the documented array size");

struct RigidConfig {
  int nr, nz, kill_width, history;
  int precision, moving_body, average_psi, implicit_diffusion;
  double dx, nu, brink_lam, U0, T_ramp, CFL, eps, dt_diffusive;
  double osc_amp, osc_omega, drag_scale, drag_div;
  double dt_cap, period, t_end;
  double body_z0, body_r0, body_aspect2, body_radius, body_disp, moll_zone;
};

struct RigidBuffers {
  void *vorticity, *psi, *u_z, *u_r, *u_pen_z, *u_pen_r, *char_func, *work0, *work1, *avg_psi;
  double* scalars;
  double* drag;
  const double* r;
  const double* z;
  const double* kill_wz;
  const double* kill_wr;
};

constexpr int kRigidStages = 7;
extern const char* const kRigidStageNames[kRigidStages];

// `batch` cases of one (nr, nz) grid stepped together; every field buffer
// holds the cases back to back, and each case has its own configuration
// (viscosity, body, ramp, dt rule, force scale) and device scalars.
struct ImplicitPlan;

struct RigidStepperBase {
  RigidConfig cfg{};                // case 0 (grid-level fields are shared)
  RigidConfig* dcfg = nullptr;      // per-case configurations, device
  int batch = 1;
  RigidBuffers buf{};
  StokesPlan* plan = nullptr;
  ImplicitPlan* implicit = nullptr;  // implicit_diffusion: the diffusion solve after step_fused
  void* spec = nullptr;        // spectral work buffer, batch * nr * nz doubles
  double* partials = nullptr;  // force partial sums of the fused blocks
  double* slots = nullptr;     // velocity maximum slots
  int blocks_x = 0, npartials = 0;
  // The step reads the stored vorticity from one buffer of the pair
  // (vorticity, work0) and writes the other; parity says which holds it.
  int parity = 0;
  // Graphs of 1, 2, 4, 8 and 16 consecutive steps, one per starting parity,
  // each also in a form that ends the run (the last step's completion and,
  // at an odd final parity, the copy home).  A run launches n / 16 graphs of
  // 16 steps and then one graph per set bit of n mod 16, the last in its
  // ending form: each graph launch leaves the device idle for several
  // microseconds before its first kernel (~8 us measured), so a run takes
  // as few launches as its step count allows.
  static constexpr int kGraphSizes = 5;  // 2^k steps, k < kGraphSizes
  static constexpr int kChunkSteps = 1 << (kGraphSizes - 1);
  hipGraph_t graphs[kGraphSizes][2][2] = {};  // [k][starting parity][ends the run]
  hipGraphExec_t graph_execs[kGraphSizes][2][2] = {};
  hipStream_t graph_stream = nullptr;
  // graph launches so far: 16-step graphs and shorter ones
  int64_t chunk_launches = 0, step_launches = 0;

  hipError_t init(const RigidConfig* cfgs, int nbatch, const RigidBuffers& b);
  virtual hipError_t enqueue_stage(int stage, hipStream_t s) = 0;
  bool stage_enabled(int stage) const;
  hipError_t enqueue_step(hipStream_t s);
  hipError_t run(int nsteps, bool use_graph, hipStream_t s);
  hipError_t capture_graphs(hipStream_t s);
  void drop_graphs();
  // Brings the current vorticity back to the home buffer (after odd counts).
  virtual hipError_t settle(hipStream_t s) = 0;
  // u_pen of the last step into the u_pen buffers (on demand).
  virtual hipError_t velocity_out(hipStream_t s) = 0;
  virtual ~RigidStepperBase();
};

template <typename T>
struct RigidStepper : RigidStepperBase {
  hipError_t enqueue_stage(int stage, hipStream_t s) override;
  hipError_t settle(hipStream_t s) override;
  hipError_t velocity_out(hipStream_t s) override;
};

// Cases of several grids in one launch per pass (axs_ragged.hpp): runs of
// cases with one (nr, nz, dx) form groups, each with its own plan; every
// field holds all cases back to back, and the r and kill-weight tables hold
// each case's table back to back.  No moving body; nz <= 1024.
template <typename T>
struct RaggedStepper : RigidStepperBase {
  std::vector<StokesPlan*> plans;
  std::vector<RaggedGroup> groups;
  RaggedGroup* dgroups = nullptr;
  int rows = 0, wave_blocks = 0, tri_blocks = 0, vt_tiles = 0, ft_tiles = 0, maxM = 0;
  int tri_L = 8;    // radial segment length (timed search over 4, 8, 16 at set-up)
  int vmax_tr = 16;  // velocity_max tile rows (timed search over 16, 8, 32 at set-up)
  size_t tri_shm = 0;
  int64_t cells = 0;
  hipError_t init_ragged(const RigidConfig* cfgs, int nbatch, const RigidBuffers& b);
  hipError_t enqueue_stage(int stage, hipStream_t s) override;
  hipError_t settle(hipStream_t s) override;
  hipError_t velocity_out(hipStream_t s) override;
  ~RaggedStepper() override;
};

}  // namespace axs
