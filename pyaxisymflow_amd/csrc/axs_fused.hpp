// Fused stencil passes and the device-side step scalars they finalize.
#pragma once

#include "axs_stepper.hpp"

namespace axs {

// Velocity maximum slots: kMaxSlots doubles, kSlotStride apart (4 KiB, so
// they spread over memory channels), holding bit patterns of non-negative doubles; all zero between
// steps (finalize_step clears them).
constexpr int kMaxSlots = 64, kSlotStride = 512;
constexpr int kSlotsPerCase = kMaxSlots * kSlotStride;  // doubles per case of a batch

// Ramp factor, penalisation velocity and body centre at time t.  With no
// oscillation (omega = 0) cos and sin of 0 are 1 and 0 exactly, so they are
// not evaluated.
__device__ __forceinline__ void next_step_scalars(double* sc, const RigidConfig& cfg, double t) {
  sc[kScPrefac] = (cfg.T_ramp > 0.0 && t < cfg.T_ramp) ? sin(0.5 * M_PI * t / cfg.T_ramp) : 1.0;
  if (cfg.osc_omega == 0.0) {
    sc[kScUpen] = cfg.osc_amp;
    sc[kScBodyZ] = cfg.body_z0;
  } else {
    sc[kScUpen] = cfg.osc_amp * cos(cfg.osc_omega * t);
    sc[kScBodyZ] = cfg.body_z0 + cfg.body_disp * sin(cfg.osc_omega * t);
  }
}

// Scalars of one step derived from max(|u_z| + |u_r|).
struct DtScalars {
  double dt, fr, reset;
};

// dt rule: flow_past_sphere.py:150-153 and spheroid_streaming.py:139-201
// (period timer reset via numpy.isclose, 3-term min taking the first
// minimal value, period and tEnd clipping).  Pure: every block that needs
// dt evaluates it from the same inputs.
__device__ __forceinline__ DtScalars compute_dt(const double* sc, double m,
                                                const RigidConfig& cfg) {
  double fr = sc[kScFreqTimer];
  double reset = 0.0;
  if (cfg.period > 0.0 && fabs(fr - cfg.period) <= 1e-8 + 1e-5 * fabs(cfg.period)) {
    fr = 0.0;
    reset = 1.0;
  }
  // implicit diffusion: the fixed step the diffusion plan was built for
  // (flow_past_sphere.py:146-148, spheroid_streaming.py:190-191)
  if (cfg.implicit_diffusion) return DtScalars{cfg.dt_diffusive, fr, reset};
  double dt = cfg.dt_diffusive;
  const double cfl_dt = cfg.CFL * cfg.dx / (m + cfg.eps);
  if (cfl_dt < dt) dt = cfl_dt;
  if (cfg.dt_cap < dt) dt = cfg.dt_cap;
  if (cfg.period > 0.0 && fr + dt > cfg.period) dt = cfg.period - fr;
  if (cfg.t_end > 0.0 && sc[kScT] + dt > cfg.t_end) dt = cfg.t_end - sc[kScT];
  return DtScalars{dt, fr, reset};
}

// Stores the step's dt-derived scalars (one thread, during the pass whose
// blocks read only the slots it does not write).
__device__ __forceinline__ void store_dt_scalars(double* sc, const DtScalars& d, double m,
                                                 const RigidConfig& cfg) {
  sc[kScFreqStep] = d.fr;
  sc[kScPrefacStep] = sc[kScPrefac];
  sc[kScUpenStep] = sc[kScUpen];
  sc[kScPending] = 1.0;
  sc[kScAvgReset] = d.reset;
  sc[kScMaxU] = m;
  sc[kScDt] = d.dt;
  sc[kScDtByDx] = d.dt / cfg.dx;
  sc[kScLamDt] = cfg.brink_lam * d.dt;
  sc[kScHalfNuDt] = 0.5 * cfg.nu * d.dt;
  sc[kScNuDt] = cfg.nu * d.dt;
}

// Force / drag coefficient into the history, then t += dt, it += 1, the
// period timer, and the next step's ramp factor, U_pen and body centre.
__device__ __forceinline__ void finalize_step_scalars(double* sc, double* drag, double s,
                                                      const RigidConfig& cfg) {
  const long it = (long)sc[kScIt];
  if (drag && it < cfg.history) drag[it] = (cfg.drag_scale * s) / cfg.drag_div;
  sc[kScDragSum] = s;
  const double t = sc[kScT] + sc[kScDt];
  sc[kScT] = t;
  sc[kScFreqTimer] = sc[kScFreqStep] + sc[kScDt];
  sc[kScIt] = double(it + 1);
  next_step_scalars(sc, cfg, t);
}


// Every launcher runs `batch` cases of one (nr, nz) grid: fields back to
// back (nr * nz each), AXS_RIGID_SCALARS scalars, kSlotsPerCase slots, one
// block's worth of partials and `history` drag entries per case, and the
// per-case configurations cfgs[batch] in device memory.
// Completes a case's last step if its pending flag is set: fixed-order sum
// of the force partials (thread t adds t, t + blockDim, ...; then a tree in
// red[blockDim]), force record, t += dt, it += 1, period timer, next step's
// scalars, cleared maximum slots.  Called by one whole block.
__device__ __forceinline__ void finalize_pending(const double* __restrict__ partials, int count,
                                                 double* __restrict__ slots,
                                                 double* __restrict__ sc,
                                                 double* __restrict__ drag,
                                                 const RigidConfig& cfg, double* red) {
  if (sc[kScPending] == 0.0) return;
  const int nt = blockDim.x;
  double s = 0.0;
  // sixteen independent loads in flight per batch; added in index order
  for (int base = threadIdx.x; base < count; base += 16 * nt) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int q = base + u * nt;
      v[u] = (q < count) ? partials[q] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) s += v[u];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < nt; off <<= 1) {
    // pairwise in a fixed pattern (works for any block size)
    double v = 0.0;
    const bool take = (threadIdx.x % (2 * off) == 0) && (threadIdx.x + off < nt);
    if (take) v = red[threadIdx.x + off];
    __syncthreads();
    if (take) red[threadIdx.x] = red[threadIdx.x] + v;
    __syncthreads();
  }
  for (int q = threadIdx.x; q < kMaxSlots; q += nt) slots[q * kSlotStride] = 0.0;
  if (threadIdx.x == 0) {
    finalize_step_scalars(sc, drag, red[0], cfg);
    sc[kScPending] = 0.0;
  }
}

template <typename T>
hipError_t launch_velocity_max(const T* psi, const double* r, int nr, int nz, T dx,
                               const double* sc, double* slots, const RigidConfig* cfgs,
                               int batch, hipStream_t s);
// Number of step_fused blocks (= force partials) for a grid.
int fused_blocks(int nr, int nz);
// Penalisation through the end of the step: reads psi, chi and the stored
// vorticity wsrc (kill taper applied on the fly; wz: 2 kw weights, wr: kw
// weights), writes the new vorticity to wdst (must not alias wsrc), the psi
// average (if avg) and one force partial per block.
template <typename T>
hipError_t launch_step_fused(const T* psi, const T* chi, const T* wsrc, T* wdst, T* avg,
                             const double* r, const double* wz, const double* wr, int kw, int nr,
                             int nz, T dx, double* sc, const double* slots, double* partials,
                             const RigidConfig* cfgs, int batch, hipStream_t s, bool rk2 = true);
template <typename T>
hipError_t launch_velocity_out(const T* psi, const T* chi, T* pz, T* pr, const double* r, int nr,
                               int nz, T dx, const double* sc, const RigidConfig* cfgs, int batch,
                               hipStream_t s);
// Uneven batches (axs_ragged.hpp): tiles of every case of every group in
// one launch; per-case scalars, slots and configurations at the global case
// index, force partials at part0 + (case - case0) * ftiles of the group.
int fused_tiles_z(int nz);
int vmax_tiles_z(int nz);
// velocity_max tiles of one case with tiles of tr rows (8, 16 or 32) by 64
int vmax_tiles(int nr, int nz, int tr);
template <typename T>
hipError_t launch_velocity_max_ragged(const T* psi, const double* sc, double* slots,
                                      const RigidConfig* cfgs, const RaggedGroup* grp, int ngrp,
                                      int tiles, int tr, hipStream_t s);
template <typename T>
hipError_t launch_step_fused_ragged(const T* psi, const T* chi, const T* wsrc, T* wdst, T* avg,
                                    int kw, double* sc, const double* slots, double* partials,
                                    const RigidConfig* cfgs, const RaggedGroup* grp, int ngrp,
                                    int tiles, hipStream_t s);
hipError_t launch_finalize_ragged(const double* partials, double* slots, double* sc, double* drag,
                                  const RigidConfig* cfgs, const RaggedGroup* grp, int ngrp,
                                  int ncases, hipStream_t s);
// One block: force sum of the partials, then t, it, period timer, the next
// step's scalars, and cleared maximum slots.
hipError_t launch_finalize_step(const double* partials, int count, double* slots, double* sc,
                                double* drag, const RigidConfig* cfgs, int batch, hipStream_t s);

}  // namespace axs
