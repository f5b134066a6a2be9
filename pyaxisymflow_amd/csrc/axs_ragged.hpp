// Uneven batches: cases of several grids advanced by one launch per stage.
//
// Cases come in runs of one grid ("groups", in the caller's order); every
// field holds the cases back to back, each case nr * nz cells.  A stage's
// launch enumerates the work items of all groups in order (rows for the
// z-transforms, mode chunks for the radial solves, tiles for the stencil
// passes), and each workgroup finds its group from the first item index of
// every group, then its case and item within the group.  Per-group tables
// (transform twiddles, radial coefficients, r, kill weights) come from the
// group's own plan.
#pragma once

#include "axs_fft.hpp"

namespace axs {

constexpr int kMaxRaggedGroups = 16;

struct RaggedGroup {
  int nr, nz, ncase, case0;  // grid, cases, global index of the first case
  int rb0, rpb, bs;          // row passes: first workgroup of the group, rows per
                             // 256-thread workgroup, threads per row (rpb bs = 256)
  int tri0, K, P, tchunks;   // radial solve: first workgroup, modes per workgroup,
                             // segments, workgroups per case
  int vt0, vz, vtiles;       // velocity_max: first tile, tiles along z, tiles per case
  int ft0, fz, ftiles;       // step_fused: first tile, tiles along z, tiles per case
  int part0;                 // first force partial of the group
  int wave;                  // 1: the row passes run one-wave register transforms
                             // (nz = 256, 384, 512, 768, 1024): no Stockham row blocks here
  int wr0;                   // first one-wave row block of the group (wave groups)
  int64_t cell0;             // first cell of the group's first case
  double dx;
  FftPlanArgs fp;
  const Cplx<double>* tw;
  const Cplx<double>* e1;
  const Cplx<double>* e2;
  const double* rvec;  // r_i scale of the forward transform
  const double* a;
  const double* b;
  const double* c;
  const double* lam;
  const double* cseg;
  const double* r;
  const double* kill_wz;
  const double* kill_wr;
};

// Index of the group whose range [g.*F, next.*F) holds item b.
template <int RaggedGroup::*F>
__device__ __forceinline__ int ragged_group(const RaggedGroup* __restrict__ g, int n, int b) {
  int q = 0;
  while (q + 1 < n && b >= g[q + 1].*F) ++q;
  return q;
}

}  // namespace axs
