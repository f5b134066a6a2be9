// Batched soft-body steps: many soft-sphere cases, each with its own grid,
// advanced by one launch per pass over all of them (the serving form of
// examples/SoftSphereInTaylorGreenVortex/soft_sphere_in_taylor_green_vortex.py:101-213).
//
// Every pass takes the device table of SoftCase entries; the case of a
// workgroup is blockIdx.z for the row-parallel passes (grid sized for the
// largest case; a workgroup beyond its own case's rows or columns exits at
// once) and blockIdx.y for the passes with one workgroup per case or per
// row.  A case's arithmetic is the single-case step's, so a case gives the
// same bits alone and beside others.
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
  kSsCount
};

struct SoftCase {
  int nr, nz, kill_width, reinit_iterations;
  double dx, nu, G, rho_f, CFL, eps, r_ball, z_cm, r_cm, moll_zone, extrap_zone, reinit_band;
  // caller-owned state
  double *vorticity, *psi, *u_z, *u_r, *eta1, *eta2, *phi, *chi, *scalars;
  const double *r, *z, *kill_wz, *kill_wr;
  // step scratch
  double *w0, *w1, *tmp, *spec, *e1n, *e2n, *slots, *e1d, *e2d;
  int16_t *cur, *tgt;
  double *s11, *s12, *s22, *g1z, *g1r, *g2z, *g2r, *tz, *tr;
  ReinitWork rw;
  LsqWork lw;
  unsigned* need;
  // row transforms (any length) and the radial solve's tables
  AnyDftArgs any;
  const Cplx<double>* ge2;
  const double *ta, *tb, *tc, *tlam, *tcseg, *rvec;
  int tK, tP;
};

// Launch shapes of a batch (host side): the largest case's dimensions and
// the any-length transform's block shape over all cases.
struct SoftBatchDims {
  int ncase = 0, max_nr = 0, max_nz = 0, max_chunks = 0;
  int any_nth = 64, any_per = 1;
  size_t any_lds = 0, tri_lds = 0;
};

// Radial segment length of every case of a batch (the tridiagonal launch
// is compiled for one L).
constexpr int kSoftBatchTriL = 8;

// Passes in other translation units (one launch each over all cases).
hipError_t launch_soft_batch_dct2(const SoftCase* cases, const SoftBatchDims& d, hipStream_t s);
hipError_t launch_soft_batch_dct3(const SoftCase* cases, const SoftBatchDims& d, hipStream_t s);
hipError_t launch_soft_batch_tridiag(const SoftCase* cases, const SoftBatchDims& d,
                                     hipStream_t s);
hipError_t launch_soft_batch_velocity(const SoftCase* cases, const SoftBatchDims& d,
                                      hipStream_t s);
hipError_t launch_soft_batch_refmap(const SoftCase* cases, const SoftBatchDims& d, hipStream_t s);
hipError_t launch_soft_batch_reinit(const SoftCase* cases, const SoftBatchDims& d, int iterations,
                                    hipStream_t s);
hipError_t launch_soft_batch_lsq_list(const SoftCase* cases, const SoftBatchDims& d,
                                      hipStream_t s);
hipError_t launch_soft_batch_lsq_sweep(const SoftCase* cases, const SoftBatchDims& d, int sweep,
                                       hipStream_t s);
hipError_t launch_soft_batch_lsq_tail(const SoftCase* cases, const SoftBatchDims& d, int sweeps,
                                      hipStream_t s);
hipError_t launch_soft_batch_stress(const SoftCase* cases, const SoftBatchDims& d, hipStream_t s);
hipError_t launch_soft_batch_advect(const SoftCase* cases, const SoftBatchDims& d, hipStream_t s);
hipError_t launch_soft_batch_diffuse(const SoftCase* cases, const SoftBatchDims& d, hipStream_t s);

}  // namespace axs
