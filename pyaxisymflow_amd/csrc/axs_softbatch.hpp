// Batched soft-body steps: many soft-sphere cases, each with its own grid,
// advanced by one launch per pass over all of them (the serving form of
// examples/SoftSphereInTaylorGreenVortex/soft_sphere_in_taylor_green_vortex.py:101-213).
//
// Every pass takes the device table of SoftCase entries; the per-cell
// passes take a block table (256 cells of one row of one case per
// workgroup), the row transforms and the radial solve a grid sized for the
// largest case (blockIdx.y = case; workgroups beyond a case's rows or mode
// chunks exit at once), and the per-case passes one workgroup per case.  A
// case's arithmetic is the single-case step's, so a case gives the same
// bits alone and beside others.
//
// A batch stores its flow fields (vorticity, psi, velocity, the
// spectrum, the stress scratch and the solid mask) in one type B: double
// for the FP64 step, float for the FP32-bulk step.  Both keep FP64 for the
// radial elimination (its arithmetic; the spectrum is stored in B), the
// reference map, the level set, its reinit, the least-squares
// extrapolation and the scalars: the interface's thresholds decide the
// mask and the extrapolation's cells, and those stay the FP64 ones.
#pragma once

#include "axs_anydft.hpp"
#include "axs_lsq.hpp"
#include "axs_reinit.hpp"

namespace axs {

// scalar slots of a soft step (AXS_RIGID_SCALARS doubles per case)
enum : int {
  kSsT = 0,           // time
  kSsDt,              // dt of the current step
  kSsDtByDx,          // dt / dx
  kSsHalfNuDt,        // 0.5 nu dt
  kSsNuDt,            // nu dt
  kSsMaxU,            // max(|u_z| + |u_r|)
  kSsIt,              // step counter
  kSsLsqSweeps,       // sweeps that fitted cells in the last step
  kSsLsqUnconverged,  // steps whose last allowed sweep still fitted cells
  kSsFreqTimer,       // streaming: time since the cycle timer last reset
  kSsUwall,           // streaming: tethered core's velocity U_0 cos(omega t)
  kSsZtether,         // streaming: tethered core's centre Z_cm + e r_ball sin(omega t)
  kSsCount
};

template <typename B>
struct SoftCaseT {
  int nr, nz, kill_width, reinit_iterations;
  double dx, nu, G, rho_f, CFL, eps, r_ball, z_cm, r_cm, moll_zone, extrap_zone, reinit_band;
  // caller-owned state
  B *vorticity, *psi, *u_z, *u_r, *chi;
  double *eta1, *eta2, *phi, *scalars;
  const double *r, *z, *kill_wz, *kill_wr;
  // step scratch
  B *w0, *w1, *tmp, *spec;
  double *e1n, *e2n, *slots, *e1d, *e2d;
  int16_t *cur, *tgt;
  B *s11, *s12, *s22, *g1z, *g1r, *g2z, *g2r, *tz, *tr;
  ReinitWork rw;
  int reinit_tile_base;  // the case's first entry in the batch's reinit tile table
  LsqWork lw;
  unsigned* need;
  // rows [zone[0], zone[1]] hold the step's extrapolation target zone,
  // [zone[2], zone[3]] the previous step's
  int* zone;
  // row transforms (any length) and the radial solve's tables
  AnyDftArgs any;
  const Cplx<double>* ge2;
  const Cplx<double>*te1, *te2;  // the plan's e1, e2 (half-length cosine rows)
  int any_half;
  const double *ta, *tb, *tc, *tlam, *tcseg, *rvec;
  int tK, tP;
};
using SoftCase = SoftCaseT<double>;

// Launch shapes of a batch (host side): the largest case's dimensions, the
// any-length transform's block shape over all cases, and the block tables
// of the per-cell passes (device arrays): one entry per 256-cell row
// segment of every case, over its (nr, nz) grid and over its doubled
// (2 nr, nz) grid, so that a per-cell pass launches exactly the workgroups
// the cases have (a grid sized for the largest case would launch mostly
// empty ones).
struct SoftBatchDims {
  int ncase = 0, max_nr = 0, max_nz = 0, max_chunks = 0;
  int any_nth = 64, any_per = 1;
  size_t any_lds = 0, tri_lds = 0;
  const uint32_t* cell_blocks = nullptr;
  int ncell_blocks = 0;
  const uint32_t* dbl_blocks = nullptr;
  int ndbl_blocks = 0;
  const uint32_t* reinit_tiles = nullptr;  // reinit tiles of every case's grid
  int nreinit_tiles = 0;
  unsigned* reinit_list = nullptr;  // two work lists of nreinit_tiles entries
  unsigned* reinit_cnt = nullptr;   // their three counters
};

// Tile of the level-set reinit's tiled sweeps (rows x columns).
constexpr int kReinitTileRows = 8, kReinitTileCols = 32;

// Block-table entry: case (10 bits), row (13 bits), 256-column block (9 bits).
constexpr int kSoftBatchMaxCases = 1024;
__host__ __device__ constexpr uint32_t soft_block_entry(int c, int row, int colblk) {
  return uint32_t(c) | (uint32_t(row) << 10) | (uint32_t(colblk) << 23);
}
struct SoftBlock {
  int c, i, j;  // case, row, this thread's column
};
__device__ __forceinline__ SoftBlock soft_block(const uint32_t* __restrict__ t) {
  const uint32_t v = t[blockIdx.x];
  return SoftBlock{int(v & 1023u), int((v >> 10) & 8191u),
                   int(v >> 23) * 256 + int(threadIdx.x)};
}

// Rows of a case that the passes after the mirror (copy back, stress,
// stress divergence, extrapolation list) visit: the extrapolation's target
// zone, recorded by the mirror pass of the same step, and kZoneMargin rows
// either side; every row in the case's first step.  Beyond them the
// reference map is zero (the copy back zeroes it outside the target zone,
// and one step's ENO3 spreads it by two rows while the zone moves by less
// than one), so the stress and its divergence are +-0 there and leave the
// vorticity's value unchanged: skipping those rows changes no value (only,
// at most, the sign of a zero in the reference map or the stress scratch).
constexpr int kZoneMargin = 6;
template <class CaseT>
__device__ __forceinline__ bool zone_row(const CaseT& C, int i) {
  if (C.scalars[kSsIt] == 0.0) return true;
  return i >= C.zone[0] - kZoneMargin && i <= C.zone[1] + kZoneMargin;
}
// The passes before the mirror (reference-map ENO3, pin, the mirror
// itself) go by the previous step's zone, two rows wider: the zone moves by
// less than a row per step.  A case's zone starts as every row, so its first
// step in a batch visits the whole grid even when its state was loaded.
template <class CaseT>
__device__ __forceinline__ bool prev_zone_row(const CaseT& C, int i) {
  if (C.scalars[kSsIt] == 0.0) return true;
  return i >= C.zone[2] - kZoneMargin - 2 && i <= C.zone[3] + kZoneMargin + 2;
}

// Radial segment length of every case of a batch (the tridiagonal launch
// is compiled for one L).
constexpr int kSoftBatchTriL = 8;

// Passes in other translation units (one launch each over all cases),
// instantiated for B = double and B = float.
template <typename B>
hipError_t launch_soft_batch_dct2(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                  hipStream_t s);
template <typename B>
hipError_t launch_soft_batch_dct3(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                  hipStream_t s);
template <typename B>
hipError_t launch_soft_batch_tridiag(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                     hipStream_t s);
template <typename B>
hipError_t launch_soft_batch_velocity(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                      hipStream_t s);
template <typename B>
hipError_t launch_soft_batch_refmap(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                    hipStream_t s);
template <typename B>
hipError_t launch_soft_batch_reinit(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                    int iterations, hipStream_t s);
template <typename B>
hipError_t launch_soft_batch_lsq_list(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                      hipStream_t s);
template <typename B>
hipError_t launch_soft_batch_lsq_sweep(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                       int sweep, hipStream_t s);
template <typename B>
hipError_t launch_soft_batch_lsq_tail(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                      int sweeps, hipStream_t s);
template <typename B>
hipError_t launch_soft_batch_stress(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                    hipStream_t s);
template <typename B>
hipError_t launch_soft_batch_advect(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                    hipStream_t s);
template <typename B>
hipError_t launch_soft_batch_diffuse(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                     hipStream_t s);

}  // namespace axs
