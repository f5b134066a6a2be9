// Implicit Euler diffusion of ImplicitEulerDiffusionStepper
// (kernels/implicit_diffusion_solver.py:6-143): solves
//   (I - nu dt L) w' = w,   L w = A_r w + w P_z^T,
// P_z = [1, -2, 1] / dx^2 with zero ghost values at both z ends (:46-52),
// A_r = P_r + D_r / r - I / r^2 (:53-70), restated as
//   1. row pass:    what = DST-I_z(w)             (S w^T, S the orthonormal DST-I)
//   2. column pass: (1 - nu dt (A_r + lambda_k)) w'hat_k = what_k per z-mode k
//   3. row pass:    w' = DST-I_z(w'hat)           (S is symmetric and S S = I)
// The Dirichlet P_z has the eigenvectors sin(pi (j+1)(k+1) / (nz+1)) and the
// eigenvalues lambda_k = -(4 / dx^2) sin^2(pi (k+1) / (2 (nz+1))).  The
// radial matrix is tridiagonal and strictly diagonally dominant (diagonal
// 1 + nu dt (2 / dx^2 + 1 / r^2 - lambda_k) against off-diagonals summing to
// nu dt 2 / dx^2), so the column pass of the Stokes solve (segmented Thomas
// elimination without pivoting) applies with other coefficient tables.
// A fast sine transform would need complex FFTs of nz + 1 points (257 is
// prime, 2049 = 3 x 683 for the power-of-two grids), so the row passes are
// the dense product with S, O(nz^2) per row, on LDS tiles.
#pragma clang fp contract(fast)

#include "axs_common.hpp"
#include "axs_implicit.hpp"

#include <cmath>
#include <cstdlib>
#include <new>
#include <vector>

namespace axs {

// out[i, k] = sum_j in[i, j] S[j, k] for rows i < rows (S symmetric, nz x
// nz).  Workgroup: a BM-row x 64-column output tile, 256 threads, each
// thread BM / 16 rows x 4 columns (rows ty + 16 h, columns tx + 16 c); j
// advances in chunks of 16 staged in LDS, two LDS buffers so the global
// loads of the next chunk are in flight while this one is multiplied (one
// barrier per chunk).  Tiles that reach past the grid read zeros there and
// skip those stores.
constexpr int kDstBN = 64, kDstBK = 16;

template <int BM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) dst1_rows_kernel(const double* __restrict__ in,
                                                         double* __restrict__ out,
                                                         const double* __restrict__ S, int rows,
                                                         int nz) {
  constexpr int RH = BM / 16;                 // rows per thread
  constexpr int NA = BM * kDstBK / 256;       // A elements per thread per chunk
  constexpr int NB = kDstBK * kDstBN / 256;   // B elements per thread per chunk
  __shared__ double sA[2][kDstBK][BM + 1];
  __shared__ double sB[2][kDstBK][kDstBN];
  const int t = threadIdx.x;
  const int tx = t & 15, ty = t >> 4;
  const int col0 = blockIdx.x * kDstBN;
  const int row0 = blockIdx.y * BM;
  double acc[RH][4] = {};
  double va[NA], vb[NB];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < NA; ++q) {  // A: BM x 16, element e = t + 256 q
      const int e = t + 256 * q, r = e >> 4, kk = e & 15;
      const int gi = row0 + r, gj = k0 + kk;
      va[q] = (gi < rows && gj < nz) ? in[(int64_t)gi * nz + gj] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {  // B: 16 x 64, element e = t + 256 q
      const int e = t + 256 * q, kk = e >> 6, cc = e & 63;
      const int gj = k0 + kk, gk = col0 + cc;
      vb[q] = (gj < nz && gk < nz) ? S[(int64_t)gj * nz + gk] : 0.0;
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int e = t + 256 * q;
      sA[buf][e & 15][e >> 4] = va[q];
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int e = t + 256 * q;
      sB[buf][e >> 6][e & 63] = vb[q];
    }
  };
  load(0);
  stage(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < nz; k0 += kDstBK) {
    const bool more = k0 + kDstBK < nz;
    if (more) load(k0 + kDstBK);
    const double(*pa)[BM + 1] = sA[buf];
    const double(*pb)[kDstBN] = sB[buf];
#pragma unroll 4
    for (int kk = 0; kk < kDstBK; ++kk) {
      double a[RH], bv[4];
#pragma unroll
      for (int h = 0; h < RH; ++h) a[h] = pa[kk][ty + 16 * h];
#pragma unroll
      for (int c = 0; c < 4; ++c) bv[c] = pb[kk][tx + 16 * c];
#pragma unroll
      for (int h = 0; h < RH; ++h)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[h][c] = fma(a[h], bv[c], acc[h][c]);
    }
    if (more) {
      // the other buffer was last read in the previous chunk, before the
      // barrier that ended it
      stage(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }
#pragma unroll
  for (int h = 0; h < RH; ++h) {
    const int gi = row0 + ty + 16 * h;
    if (gi >= rows) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int gk = col0 + tx + 16 * c;
      if (gk < nz) out[(int64_t)gi * nz + gk] = acc[h][c];
    }
  }
}

// Tile height: 64 rows when that still gives every CU a workgroup, else 32
// (AXS_DST_BM = 32 or 64 pins it).
static int dst_tile_rows(int rows, int nz) {
  static const int pin = [] {
    const char* v = std::getenv("AXS_DST_BM");
    return v ? std::atoi(v) : 0;
  }();
  if (pin == 32 || pin == 64) return pin;
  return (int64_t)ceil_div(nz, kDstBN) * ceil_div(rows, 64) >= 256 ? 64 : 32;
}

static hipError_t launch_dst1_rows(const double* in, double* out, const double* S, int rows,
                                   int nz, hipStream_t s) {
  if (dst_tile_rows(rows, nz) == 64)
    dst1_rows_kernel<64><<<dim3(ceil_div(nz, kDstBN), ceil_div(rows, 64)), 256, 0, s>>>(
        in, out, S, rows, nz);
  else
    dst1_rows_kernel<32><<<dim3(ceil_div(nz, kDstBN), ceil_div(rows, 32)), 256, 0, s>>>(
        in, out, S, rows, nz);
  return hipGetLastError();
}

ImplicitPlan::~ImplicitPlan() {
  if (S) (void)hipFree(S);
}

hipError_t ImplicitPlan::init(int nr, int nz, double dx, double nu_dt_, int batch) {
  nu_dt = nu_dt_;
  tri.nr = nr;
  tri.nz = nz;
  tri.dx = dx;
  tri.tune_key = 1;
  // radial coordinate as implicit_diffusion_solver.py:22-26 (numpy.linspace)
  std::vector<double> r(nr);
  const double start = dx / 2, stop = nr * dx - dx / 2;
  const double step = (stop - start) / (nr - 1);
  for (int i = 0; i < nr; ++i) r[i] = i * step + start;
  r[nr - 1] = stop;
  // A_r entries as :44-70 form them: inv_dx2 * {1, -2, 1} plus
  // inv_2dx * {-1, 1} / r on the off-diagonals and -1 / r^2 on the diagonal
  const double inv_dx2 = 1.0 / dx / dx, inv_2dx = 1.0 / 2.0 / dx;
  std::vector<double> a(nr, 0.0), b(nr), c(nr, 0.0), lam(nz);
  for (int i = 0; i < nr; ++i) {
    const double sub = inv_dx2 * 1.0 + (inv_2dx * -1.0) / r[i];
    const double sup = inv_dx2 * 1.0 + (inv_2dx * 1.0) / r[i];
    const double dia = inv_dx2 * -2.0 - 1.0 / (r[i] * r[i]);
    b[i] = 1.0 - nu_dt * dia;
    if (i > 0) a[i] = -nu_dt * sub;
    if (i < nr - 1) c[i] = -nu_dt * sup;
  }
  for (int k = 0; k < nz; ++k) {
    const double sn = std::sin(M_PI * (double)(k + 1) / (2.0 * (nz + 1)));
    lam[k] = -nu_dt * (-(4.0 / (dx * dx)) * sn * sn);
  }
  // S[j, k] = sqrt(2 / (nz + 1)) sin(pi (j+1)(k+1) / (nz + 1)), the angle
  // reduced modulo 2 pi in integers first
  std::vector<double> h((size_t)nz * nz);
  const double norm = std::sqrt(2.0 / (nz + 1));
  const int64_t period = 2 * (int64_t)(nz + 1);
  for (int j = 0; j < nz; ++j)
    for (int k = 0; k < nz; ++k) {
      const int64_t m = ((int64_t)(j + 1) * (k + 1)) % period;
      h[(size_t)j * nz + k] = norm * std::sin(M_PI * (double)m / (double)(nz + 1));
    }
  hipError_t e = tri.set_tridiag(a, b, c, lam);
  if (e == hipSuccess && batch > 1) {  // the spectral buffer of a whole batch
    (void)hipFree(tri.spec);
    tri.spec = nullptr;
    e = hipMalloc(reinterpret_cast<void**>(&tri.spec), sizeof(double) * (size_t)nr * nz * batch);
  }
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&S), sizeof(double) * h.size());
  if (e == hipSuccess) e = hipMemcpy(S, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice);
  return e;
}

hipError_t ImplicitPlan::launch(double* omega, int batch, hipStream_t s) {
  const int rows = tri.nr * batch;
  hipError_t e = launch_dst1_rows(omega, tri.spec, S, rows, tri.nz, s);
  if (e == hipSuccess) e = launch_tridiag<double>(tri, tri.spec, s, batch);
  if (e == hipSuccess) e = launch_dst1_rows(tri.spec, omega, S, rows, tri.nz, s);
  return e;
}

}  // namespace axs

using namespace axs;

struct axs_implicit_plan {
  ImplicitPlan impl;
};

extern "C" {

axs_status axs_implicit_diffusion_plan_create(int nr, int nz, double dx, double nu_dt,
                                              axs_implicit_plan** plan_out) {
  AXS_CHECK_ARG(plan_out, "null plan_out");
  *plan_out = nullptr;
  AXS_CHECK_ARG(nr >= 3 && nz >= 2, "need nr >= 3 and nz >= 2");
  AXS_CHECK_ARG(nz <= 4096 && nr <= 4096, "grid too large for the row/column passes");
  AXS_CHECK_ARG(dx > 0.0, "dx must be positive");
  AXS_CHECK_ARG(nu_dt >= 0.0 && std::isfinite(nu_dt), "nu * dt must be finite and >= 0");
  axs_implicit_plan* p = new (std::nothrow) axs_implicit_plan();
  if (!p) {
    set_error("out of host memory");
    return AXS_ERR_ALLOC;
  }
  hipError_t e = p->impl.init(nr, nz, dx, nu_dt);
  if (e != hipSuccess) {
    set_error(std::string("implicit diffusion plan: ") + hipGetErrorString(e));
    delete p;
    return AXS_ERR_HIP;
  }
  *plan_out = p;
  return AXS_OK;
}

axs_status axs_implicit_diffusion_plan_destroy(axs_implicit_plan* plan) {
  delete plan;
  return AXS_OK;
}

axs_status axs_implicit_diffusion_step_f64(axs_implicit_plan* plan, double* omega,
                                           void* stream) {
  AXS_CHECK_ARG(plan && omega, "null pointer");
  AXS_HIP(plan->impl.launch(omega, 1, as_stream(stream)));
  return AXS_OK;
}

}  // extern "C"
