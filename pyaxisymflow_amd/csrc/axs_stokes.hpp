// Stokes stream-function plan shared by the drop-in solver and the fused
// stepper.
#pragma once

#include "axs_anydft.hpp"
#include "axs_fft.hpp"
#include "axs_ragged.hpp"

#include <vector>

namespace axs {

// Length of one tridiagonal coefficient array in LDS: nr rows padded by one
// double per L rows.
__host__ __device__ constexpr int tri_coef_stride(int nr, int L) { return nr + nr / L + 1; }
// LDS bytes of one tridiagonal workgroup: four scan buffers of K P doubles,
// plus the three coefficient arrays unless they are read from global memory.
inline size_t tri_lds_bytes(int nr, int L, int K, int P, bool coef_global) {
  return sizeof(double) * (4 * (size_t)K * P + (coef_global ? 0 : 3 * (size_t)tri_coef_stride(nr, L)));
}

struct StokesPlan {
  int nr = 0, nz = 0;
  double dx = 0.0;
  FftPlanArgs fft{};
  // general: nz odd or nz / 2 with a prime factor above 5; the row passes
  // then run the any-length transforms (axs_anysize.hip) with `any`: length
  // nz / 2 for the cosine rows of an even nz (any_half; tables e1, e2),
  // else length nz with ge2[k] = e^{-i pi k / (2 nz)}, k < nz
  bool general = false;
  bool any_half = false;
  AnyDft* any = nullptr;
  Cplx<double>* ge2 = nullptr;
  // device tables
  Cplx<double>* tw = nullptr;  // e^{-2 pi i t/M}, t < M
  Cplx<double>* e1 = nullptr;  // e^{-2 pi i k/nz}, k <= M
  Cplx<double>* e2 = nullptr;  // e^{-i pi k/(2 nz)}, k <= M
  double* a = nullptr;         // sub-diagonal of A_r (a[0] = 0)
  double* b = nullptr;         // diagonal of A_r
  double* c = nullptr;         // super-diagonal of A_r (c[nr-1] = 0)
  double* lam = nullptr;       // z eigenvalues, nz
  double* rvec = nullptr;      // radial coordinate r_i, nr
  double* cseg = nullptr;      // c'_{pL-1} per segment p and mode k, (P, nz)
  double* spec = nullptr;      // spectral work buffer (nr, nz)
  int tri_L = 16, tri_P = 1, tri_K = 1;
  bool tri_G = false;  // tridiagonal coefficients read from global memory, not LDS
  // > 0: segments of force_L rows and as many modes per workgroup as fit
  // 1024 threads, no timed search (the layout an uneven batch launches)
  int force_L = 0;
  // true: the any-length row transforms for every nz (the form a soft batch
  // launches over cases of different lengths)
  bool force_general = false;
  // separates the cached column-pass searches of plans whose matrices
  // differ on the same grid (0: Stokes, 1: implicit diffusion)
  int tune_key = 0;

  std::vector<double> host_r;  // r_i on the host (init_transforms)

  // boundary conditions (AXS_BC_*): 0 Neumann z and r (DCT rows), 1 Neumann
  // r and periodic z, 2 Dirichlet r and periodic z (real-DFT rows)
  int bc = 0;

  // Transform tables and r, then the Stokes column-pass tables.
  hipError_t init(int nr, int nz, double dx, int bc = 0);
  // The DCT tables (tw, e1, e2) and r only (fft must be set).
  hipError_t init_transforms(int nr, int nz, double dx);
  // Column-pass tables (a, b, c, lam, pivots at segment starts), the
  // spectral buffer and the segment layout for the given matrices; nr and
  // nz must be set.
  hipError_t set_tridiag(const std::vector<double>& a, const std::vector<double>& b,
                         const std::vector<double>& c, const std::vector<double>& lam);
  hipError_t tune_tridiag(const std::vector<double>& a, const std::vector<double>& b,
                          const std::vector<double>& c, const std::vector<double>& lam);
  int tri_threads() const;
  ~StokesPlan();
};

struct RigidConfig;

// Extra work of the first pass of a rigid step: one more workgroup per case
// completes the previous step (force record, t, it, next scalars) while the
// row transforms run (finalize_pending in axs_fused.hpp).
struct StepFinalize {
  const double* partials = nullptr;  // null: no extra workgroup
  int count = 0;                     // partials per case
  double* slots = nullptr;
  double* sc = nullptr;
  double* drag = nullptr;
  const RigidConfig* cfgs = nullptr;
};

template <typename T>
hipError_t launch_dct2_rows(const StokesPlan& pl, const T* src, T* dst, const double* rscale,
                            int rows, hipStream_t s);
// Kill taper + DCT-II (the first pass of a fused step): src untouched,
// tapered field to `tapered` (if not null), spectrum of r * tapered to `dst`.
// `batch` cases of (nr, nz) lie back to back in every buffer.
template <typename T>
hipError_t launch_taper_dct2_rows(const StokesPlan& pl, const T* src, T* tapered, T* dst,
                                  const double* wz, const double* wr, int w, hipStream_t s,
                                  int batch = 1, const StepFinalize* fin = nullptr);
template <typename T>
hipError_t launch_dct3_rows(const StokesPlan& pl, const T* src, T* dst, int rows, hipStream_t s,
                            int batch = 1);
template <typename T>
hipError_t launch_tridiag(const StokesPlan& pl, T* d, hipStream_t s, int batch = 1);
// c'_{pL-1} of the forward elimination at every segment start p L, per mode,
// laid out (P, nz).
std::vector<double> pivot_table(const std::vector<double>& a, const std::vector<double>& b,
                                const std::vector<double>& c, const std::vector<double>& lam,
                                int nr, int nz, int L);
// Real DFT rows of a periodic-z plan (packed real spectrum, see
// rfft_rows_kernel), or their inverse.
hipError_t launch_rfft_rows(const StokesPlan& pl, const double* src, double* dst,
                            const double* rscale, int rows, hipStream_t s, bool inverse);
// One tridiagonal system a_i x_{i-1} + (b_i + lam[0]) x_i + c_i x_{i+1} = d_i
// of nr <= 4096 contiguous rows, in place; cseg = pivot_table(..., nz = 1, L = 16).
hipError_t launch_tridiag_one(double* d, const double* a, const double* b, const double* c,
                              const double* lam, const double* cseg, int nr, hipStream_t s);
// Row passes of a general plan (any nz; see StokesPlan::general).
template <typename T>
hipError_t launch_any_dct2_rows(const StokesPlan& pl, const T* src, T* tapered, T* dst,
                                const double* rscale, const double* wz, const double* wr, int w,
                                bool taper, int batch, const StepFinalize* fin, hipStream_t s);
template <typename T>
hipError_t launch_any_dct3_rows(const StokesPlan& pl, const T* src, T* dst, int rows, int batch,
                                hipStream_t s);
hipError_t launch_any_rfft_rows(const StokesPlan& pl, const double* src, double* dst,
                                const double* rscale, int rows, bool inverse, hipStream_t s);
// Orthonormal DST-I of `rows` rows of nz (D of length nz + 1, em[k] =
// e^{-i pi k / (nz + 1)}, k <= nz); src and dst must not alias.
hipError_t launch_any_dst1_rows(const AnyDft& D, const Cplx<double>* em, const double* src,
                                double* dst, int rows, int nz, hipStream_t s);
// Largest nz a plan takes: the half-length transforms cover nz <= 8192 with
// nz / 2 2,3,5-smooth; the any-length ones need their FFT length <= kAnyMaxL.
bool stokes_nz_supported(int nz);
// True when rows of half-length M run a one-wave register transform
// (M = 128, 192, 256, 384, 512, 1024; M = 2048 runs two waves per row).
bool wave_rows(int M);

// Uneven batches (axs_ragged.hpp): one launch per pass over every group.
// Row passes run 256-thread workgroups of rpb rows with bs threads each
// (nz <= 1024; groups marked wave have no row blocks here), lds_cplx complex LDS entries per workgroup (the largest
// (rpb + 1) M: the rows, then the group's twiddle table); the forward pass puts one completion workgroup per case in front
// of the rows when fin is given (force partials of case q at
// part0 + (q - case0) * ftiles of its group).
template <typename T>
hipError_t launch_taper_dct2_ragged(const T* src, T* dst, const RaggedGroup* grp, int ngrp,
                                    int ncases, int row_blocks, int w, int lds_cplx,
                                    const StepFinalize* fin, hipStream_t s);
template <typename T>
hipError_t launch_dct3_ragged(const T* src, T* dst, const RaggedGroup* grp, int ngrp,
                              int row_blocks, int lds_cplx, hipStream_t s);
// The row passes of the groups marked wave (nz = 256 to 1024) in one launch
// of wave_rows one-wavefront blocks (block b: row b - wr0 of its group):
// forward = tapered DCT-II (src field -> dst spectrum), else DCT-III.
template <typename T>
hipError_t launch_wave_rows_ragged(const T* src, T* dst, const RaggedGroup* grp, int ngrp,
                                   int wave_rows, int w, bool forward, hipStream_t s);
// Radial solves with segments of L rows (4, 8 or 16; the groups' plans
// must have been made for that L), K and P per group, one 1024-thread
// workgroup per K modes of a case.
template <typename T>
hipError_t launch_tridiag_ragged(T* d, const RaggedGroup* grp, int ngrp, int blocks, size_t shm,
                                 int L, hipStream_t s);

}  // namespace axs
