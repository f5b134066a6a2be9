// Row transforms for any nz: the DCT-II / DCT-III pair of the Neumann
// Stokes and potential plans, the real DFT pair of the periodic-z plans, and
// the DST-I of the implicit diffusion plan, each one row per workgroup with
// the row's DFT in LDS (axs_anydft.hpp: Stockham when the length factors
// into 2, 3 and 5, Bluestein's chirp-z form otherwise).  The grids of the
// drivers (nz = 256 .. 4096, nz / 2 a power of two) keep the faster
// half-length transforms of axs_stokes.hip; these cover every other nz the
// reference's dense eig route accepts (kernels/FastDiagonalisationStokesSolver.py:41-128,
// kernels/implicit_diffusion_solver.py:74-106), up to the LDS bound.
//
// DCT-II of x[0..N) by one length-N DFT (J. Makhoul, "A fast cosine
// transform in one and two dimensions", IEEE Trans. ASSP 28(1), 1980):
//   v_n = x_{2n}, v_{N-1-n} = x_{2n+1},  V = DFT_N(v),
//   X_k = Re(e^{-i pi k / (2N)} V_k)      (X_k = sum_n x_n cos(pi k (2n+1) / (2N)))
// and its inverse V_k = e^{i pi k / (2N)} (X_k - i X_{N-k}) (X_N = 0),
// v = DFT_N^{-1}(V) / N, x unshuffled; valid for odd N too.  For even N the
// real sequence v is packed as z_m = v_{2m} + i v_{2m+1} and V comes from
// one length-N/2 DFT of z (the split of dct2_post_at / dct3_pre_at, as the
// half-length transforms of the smooth lengths do): half the DFT length,
// half the LDS.
// DST-I of x[0..N) (S_k = sum_j x_j sin(pi (j+1)(k+1) / (N+1))) through the
// odd extension y of length 2M, M = N + 1 (y_0 = y_M = 0, y_m = x_{m-1},
// y_{2M-m} = -x_{m-1}), packed as z_p = y_{2p} + i y_{2p+1}:
//   Z = DFT_M(z),  Y_k = (Z_k + conj Z_{M-k}) / 2 - i e^{-i pi k / M} (Z_k - conj Z_{M-k}) / 2,
//   S_{k-1} = -Im(Y_k) / 2.
#pragma clang fp contract(fast)

#include "axs_anydft.hpp"
#include "axs_fused.hpp"
#include "axs_softbatch.hpp"
#include "axs_stokes.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

namespace axs {

// ------------------------------------------------------------------ plan

static bool radices_of(int M, FftPlanArgs* p) {
  p->M = M;
  p->nstages = 0;
  int m = M;
  for (int r : {4, 2, 3, 5}) {
    while (m % r == 0) {
      if (p->nstages >= kMaxFftStages) return false;
      p->radix[p->nstages++] = r;
      m /= r;
    }
  }
  return m == 1;
}

int any_dft_length(int N, bool* bluestein) {
  FftPlanArgs p;
  if (radices_of(N, &p)) {
    if (bluestein) *bluestein = false;
    return N;
  }
  if (bluestein) *bluestein = true;
  int L = 2 * N - 1;
  while (!radices_of(L, &p)) ++L;
  return L;
}

template <typename V>
static hipError_t upload_table(V** dev, const std::vector<V>& host) {
  hipError_t e = hipMalloc(reinterpret_cast<void**>(dev), sizeof(V) * host.size());
  if (e != hipSuccess) return e;
  return hipMemcpy(*dev, host.data(), sizeof(V) * host.size(), hipMemcpyHostToDevice);
}

// bhat = FFT_L(b) / L in place, one workgroup
template <int PER>
__global__ void __launch_bounds__(1024) any_bhat_kernel(Cplx<double>* __restrict__ b,
                                                        AnyDftArgs a) {
  extern __shared__ unsigned char smem_raw[];
  Cplx<double>* work = reinterpret_cast<Cplx<double>*>(smem_raw);
  for (int k = threadIdx.x; k < a.L; k += blockDim.x) work[k] = b[k];
  any_dft_stage(work, a, threadIdx.x, blockDim.x);
  __syncthreads();
  any_fft<double, PER>(work, a, -1, threadIdx.x, blockDim.x);
  const double inv = 1.0 / (double)a.L;
  for (int k = threadIdx.x; k < a.L; k += blockDim.x)
    b[k] = Cplx<double>{work[k].x * inv, work[k].y * inv};
}

static void allow_large_lds();
// bytes of dynamic LDS every any-kernel may take (raised by allow_large_lds)
static int g_any_lds_limit = 65536;

hipError_t AnyDft::init(int N) {
  allow_large_lds();
  bool blue = false;
  const int L = any_dft_length(N, &blue);
  if (L > kAnyMaxL || sizeof(Cplx<double>) * (size_t)L > (size_t)g_any_lds_limit)
    return hipErrorInvalidValue;
  args.N = N;
  args.L = L;
  args.bluestein = blue ? 1 : 0;
  // staged only while the doubled footprint still lets two workgroups share
  // a CU (longer rows lose more to occupancy than the LDS twiddles save)
  args.stage = 2 * sizeof(Cplx<double>) * (size_t)L <= 65536 ? 1 : 0;
  radices_of(L, &args.fp);
  nth = std::min(1024, std::max(64, ((L / 2 + 63) / 64) * 64));
  if (nth == 1024) {
    // long rows: the threads of the rows the LDS holds on one CU fill it
    // (2048 threads per CU); 1024 per row would leave room for two rows
    // where the LDS holds four (L = 2160, the 1024-point DST-I's chirp-z:
    // 90 -> 77 us per implicit step at 1024 x 512)
    const size_t lds = sizeof(Cplx<double>) * (size_t)L * (args.stage ? 2 : 1);
    const int rows = std::max(1, (int)((size_t)160 * 1024 / lds));
    nth = std::max(64, std::min(1024, (2048 / rows) / 64 * 64));
  }
  per = (L / 2 + nth - 1) / nth;
  if (per > 4) return hipErrorInvalidValue;
  per = per <= 1 ? 1 : (per <= 2 ? 2 : 4);
  std::vector<Cplx<double>> tw(L);
  for (int t = 0; t < L; ++t) {
    const double ang = -2.0 * M_PI * (double)t / (double)L;
    tw[t] = {std::cos(ang), std::sin(ang)};
  }
  Cplx<double>* d_tw = nullptr;
  hipError_t e = upload_table(&d_tw, tw);
  if (e != hipSuccess) return e;
  owned.push_back(d_tw);
  args.tw = d_tw;
  if (!blue) return hipSuccess;
  // w_n = e^{-i pi (n^2 mod 2N) / N}
  std::vector<Cplx<double>> chirp(N), b(L, Cplx<double>{0.0, 0.0});
  const int64_t two_n = 2 * (int64_t)N;
  for (int n = 0; n < N; ++n) {
    const int64_t m = ((int64_t)n * n) % two_n;
    const double ang = -M_PI * (double)m / (double)N;
    chirp[n] = {std::cos(ang), std::sin(ang)};
  }
  b[0] = {chirp[0].x, -chirp[0].y};
  for (int m = 1; m < N; ++m) {
    b[m] = {chirp[m].x, -chirp[m].y};
    b[L - m] = b[m];
  }
  Cplx<double>*d_chirp = nullptr, *d_bhat = nullptr;
  if ((e = upload_table(&d_chirp, chirp)) != hipSuccess) return e;
  owned.push_back(d_chirp);
  if ((e = upload_table(&d_bhat, b)) != hipSuccess) return e;
  owned.push_back(d_bhat);
  const size_t shm = lds_bytes<double>();
  if (per == 1)
    any_bhat_kernel<1><<<1, nth, shm, 0>>>(d_bhat, args);
  else if (per == 2)
    any_bhat_kernel<2><<<1, nth, shm, 0>>>(d_bhat, args);
  else
    any_bhat_kernel<4><<<1, nth, shm, 0>>>(d_bhat, args);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipDeviceSynchronize()) != hipSuccess) return e;
  args.chirp = d_chirp;
  args.bhat = d_bhat;
  return hipSuccess;
}

AnyDft::~AnyDft() {
  for (void* p : owned) (void)hipFree(p);
}

// ------------------------------------------------------------------ kernels

// Forward DCT-II of row i of one field (src, tapered, dst at the case):
// x = rscale_i * src row (with the kill taper of
// kill_boundary_vorticity_sine.py:4-27 applied on the fly when TAPER, as
// taper_dct2_row does; the tapered row also goes to `tapered` when not
// null), dst row = X.  Threads tid of nth, work of a.L entries in LDS; every
// thread of the block calls it (block barriers).
// The tables of one plan's DCT rows: a = the DFT (length nz, or nz / 2
// when half), ge2[k] = e^{-i pi k / (2 nz)} for k < nz (full form), e1[k] =
// e^{-2 pi i k / nz} and e2[k] = e^{-i pi k / (2 nz)} for k <= nz / 2 (half
// form).
struct AnyDctTables {
  AnyDftArgs a;
  const Cplx<double>* ge2;
  const Cplx<double>* e1;
  const Cplx<double>* e2;
  int half;
};

template <typename T, int PER, bool TAPER>
__device__ void any_dct2_row(const T* __restrict__ src, T* __restrict__ tapered,
                             T* __restrict__ dst, const double* __restrict__ rscale,
                             const double* __restrict__ wz, const double* __restrict__ wr, int nr,
                             int nz, int w, const AnyDctTables& tb, int i, Cplx<T>* work, int tid,
                             int nth) {
  const int srow = (TAPER && i >= nr - w) ? nr - w : i;
  const T* row = src + (int64_t)srow * nz;
  const T rs = rscale ? T(rscale[i]) : T(1);
  T left = T(0), right = T(0), rowf = T(1);
  if constexpr (TAPER) {
    left = row[w - 1];
    right = row[nz - w];
    rowf = (i == 0) ? T(0) : ((i >= nr - w) ? T(wr[i - (nr - w)]) : T(1));
  }
  T* out_t = (TAPER && tapered) ? tapered + (int64_t)i * nz : nullptr;
  any_dft_stage(work, tb.a, tid, nth);
  for (int j = tid; j < nz; j += nth) {
    T t = row[j];
    if constexpr (TAPER) {
      if (j < w)
        t = T(wz[j]) * left;
      else if (j >= nz - w)
        t = T(wz[w + (j - (nz - w))]) * right;
      if (i >= nr - w) t = rowf * t;
      if (i == 0) t = T(0);
      if (out_t) out_t[j] = t;
    }
    if (tb.half)
      packed_component(work, dct_dst_index(j, nz)) = rs * t;
    else
      work[dct_dst_index(j, nz)] = Cplx<T>{rs * t, T(0)};
  }
  __syncthreads();
  any_dft<T, PER>(work, tb.a, -1, tid, nth);
  T* out = dst + (int64_t)i * nz;
  if (tb.half) {
    const int M = nz / 2;
    for (int k = tid; k <= M / 2; k += nth)
      dct2_post_at<T>(work, M, k, tb.e1[k], tb.e1[M - k], tb.e2[k], tb.e2[M - k],
                      [&](int q, T val) { out[q] = val; });
    return;
  }
  for (int k = tid; k < nz; k += nth) {
    const Cplx<double> e = tb.ge2[k];
    const Cplx<T> v = work[k];
    out[k] = T(e.x) * v.x - T(e.y) * v.y;
  }
}

// Inverse (DCT-III scaled by 1 / N) of row i of one field.
template <typename T, int PER>
__device__ void any_dct3_row(const T* __restrict__ src, T* __restrict__ dst, int nz,
                             const AnyDctTables& tb, int i, Cplx<T>* work, int tid, int nth) {
  const T* row = src + (int64_t)i * nz;
  const T invN = T(1) / T(nz);
  T* out = dst + (int64_t)i * nz;
  any_dft_stage(work, tb.a, tid, nth);
  if (tb.half) {
    const int M = nz / 2;
    for (int k = tid; k < M; k += nth) {
      const int mk = M - k;  // in [1, M]
      const T xk = row[k], xnk = (k == 0) ? T(0) : row[nz - k];
      const T xm = row[mk], xnm = row[nz - mk];
      work[k] = dct3_pre_at<T>(xk, xnk, xm, xnm, tb.e2[k], tb.e2[mk], tb.e1[k]);
    }
    __syncthreads();
    any_dft<T, PER>(work, tb.a, +1, tid, nth);
    for (int j = tid; j < nz; j += nth)
      out[j] = packed_component(work, dct_dst_index(j, nz)) * invN;
    return;
  }
  const Cplx<double>* __restrict__ e2 = tb.ge2;
  for (int k = tid; k < nz; k += nth) {
    const T xk = row[k];
    const T xn = (k == 0) ? T(0) : row[nz - k];
    const Cplx<double> e = e2[k];
    work[k] = cmul(Cplx<T>{T(e.x), T(-e.y)}, Cplx<T>{xk, -xn});
  }
  __syncthreads();
  any_dft<T, PER>(work, tb.a, +1, tid, nth);
  for (int j = tid; j < nz; j += nth) out[j] = work[dct_dst_index(j, nz)].x * invN;
}

// Forward DCT-II rows; case blockIdx.y of a batch; workgroup nr completes the
// previous step when fin is given.
template <typename T, int PER, bool TAPER>
__global__ void __launch_bounds__(1024) any_dct2_rows_kernel(
    const T* __restrict__ src, T* __restrict__ tapered, T* __restrict__ dst,
    const double* __restrict__ rscale, const double* __restrict__ wz,
    const double* __restrict__ wr, int nr, int nz, int w, AnyDctTables tb, StepFinalize fin) {
  if (fin.partials && blockIdx.x == nr) {  // completes the previous step
    __shared__ double fin_red[1024];  // one entry per thread (finalize_pending)
    const int cb = blockIdx.y;
    const RigidConfig& cfg = fin.cfgs[cb];
    finalize_pending(fin.partials + (int64_t)cb * fin.count, fin.count,
                     fin.slots + (int64_t)cb * kSlotsPerCase, fin.sc + (int64_t)cb * AXS_RIGID_SCALARS,
                     fin.drag ? fin.drag + (int64_t)cb * cfg.history : nullptr, cfg, fin_red);
    return;
  }
  extern __shared__ unsigned char smem_raw[];
  const int64_t cs = (int64_t)blockIdx.y * nr * nz;
  any_dct2_row<T, PER, TAPER>(src + cs, (TAPER && tapered) ? tapered + cs : nullptr, dst + cs,
                              rscale, wz, wr, nr, nz, w, tb, blockIdx.x,
                              reinterpret_cast<Cplx<T>*>(smem_raw), threadIdx.x, blockDim.x);
}

template <typename T, int PER>
__global__ void __launch_bounds__(1024) any_dct3_rows_kernel(
    const T* __restrict__ src, T* __restrict__ dst, int nr, int nz, AnyDctTables tb) {
  extern __shared__ unsigned char smem_raw[];
  const int64_t cs = (int64_t)blockIdx.y * nr * nz;
  any_dct3_row<T, PER>(src + cs, dst + cs, nz, tb, blockIdx.x,
                       reinterpret_cast<Cplx<T>*>(smem_raw), threadIdx.x, blockDim.x);
}

// Soft batches: row blockIdx.x of case blockIdx.y (tapered omega -> w0,
// r omega -> spectrum; the spectrum -> psi), in the batch's storage type.
template <typename B, int PER>
__global__ void __launch_bounds__(1024) soft_batch_dct2_kernel(
    const SoftCaseT<B>* __restrict__ cs) {
  const SoftCaseT<B>& C = cs[blockIdx.y];
  if ((int)blockIdx.x >= C.nr) return;
  extern __shared__ unsigned char smem_raw[];
  const AnyDctTables tb{C.any, C.ge2, C.te1, C.te2, C.any_half};
  any_dct2_row<B, PER, true>(C.vorticity, C.w0, C.spec, C.rvec, C.kill_wz, C.kill_wr, C.nr,
                             C.nz, C.kill_width, tb, blockIdx.x,
                             reinterpret_cast<Cplx<B>*>(smem_raw), threadIdx.x, blockDim.x);
}

template <typename B, int PER>
__global__ void __launch_bounds__(1024) soft_batch_dct3_kernel(
    const SoftCaseT<B>* __restrict__ cs) {
  const SoftCaseT<B>& C = cs[blockIdx.y];
  if ((int)blockIdx.x >= C.nr) return;
  extern __shared__ unsigned char smem_raw[];
  const AnyDctTables tb{C.any, C.ge2, C.te1, C.te2, C.any_half};
  any_dct3_row<B, PER>(C.spec, C.psi, C.nz, tb, blockIdx.x, reinterpret_cast<Cplx<B>*>(smem_raw),
                       threadIdx.x, blockDim.x);
}

// Real DFT rows of any length into the packed real layout of the periodic
// plans: [X_0, Re X_1, Im X_1, ..., Re X_h, Im X_h (, X_{N/2} when N is
// even)], h = (N - 1) / 2; every entry one real unknown of its mode.
template <int PER>
__global__ void __launch_bounds__(1024) any_rfft_rows_kernel(
    const double* __restrict__ src, double* __restrict__ dst, const double* __restrict__ rscale,
    int nz, AnyDftArgs a) {
  extern __shared__ unsigned char smem_raw[];
  Cplx<double>* work = reinterpret_cast<Cplx<double>*>(smem_raw);
  const int i = blockIdx.x, tid = threadIdx.x, nth = blockDim.x;
  const double* row = src + (int64_t)i * nz;
  const double sc = rscale ? rscale[i] : 1.0;
  any_dft_stage(work, a, tid, nth);
  for (int j = tid; j < nz; j += nth) work[j] = Cplx<double>{sc * row[j], 0.0};
  __syncthreads();
  any_dft<double, PER>(work, a, -1, tid, nth);
  double* out = dst + (int64_t)i * nz;
  for (int p = tid; p < nz; p += nth) {
    double v;
    if (p == 0)
      v = work[0].x;
    else if (p == nz - 1 && nz % 2 == 0)
      v = work[nz / 2].x;
    else
      v = (p & 1) ? work[(p + 1) / 2].x : work[p / 2].y;
    out[p] = v;
  }
}

template <int PER>
__global__ void __launch_bounds__(1024) any_irfft_rows_kernel(const double* __restrict__ src,
                                                              double* __restrict__ dst, int nz,
                                                              AnyDftArgs a) {
  extern __shared__ unsigned char smem_raw[];
  Cplx<double>* work = reinterpret_cast<Cplx<double>*>(smem_raw);
  const int i = blockIdx.x, tid = threadIdx.x, nth = blockDim.x;
  const double* row = src + (int64_t)i * nz;
  any_dft_stage(work, a, tid, nth);
  for (int k = tid; k < nz; k += nth) {
    Cplx<double> v;
    if (k == 0) {
      v = {row[0], 0.0};
    } else if (2 * k == nz) {
      v = {row[nz - 1], 0.0};
    } else if (2 * k < nz) {
      v = {row[2 * k - 1], row[2 * k]};
    } else {  // conjugate of mode nz - k
      const int m = nz - k;
      v = {row[2 * m - 1], -row[2 * m]};
    }
    work[k] = v;
  }
  __syncthreads();
  any_dft<double, PER>(work, a, +1, tid, nth);
  const double invN = 1.0 / (double)nz;
  double* out = dst + (int64_t)i * nz;
  for (int j = tid; j < nz; j += nth) out[j] = work[j].x * invN;
}

// Real component m of the packed sequence z_p = y_{2p} + i y_{2p+1}.
__device__ __forceinline__ double& packed_y(Cplx<double>* z, int m) {
  return (m & 1) ? z[m >> 1].y : z[m >> 1].x;
}

// Orthonormal DST-I rows (sqrt(2 / (N + 1)) S), its own inverse; a has
// length M = nz + 1, em[k] = e^{-i pi k / M} for k <= nz.
template <int PER>
__global__ void __launch_bounds__(1024) any_dst1_rows_kernel(
    const double* __restrict__ src, double* __restrict__ dst, int nz, AnyDftArgs a,
    const Cplx<double>* __restrict__ em, double norm) {
  extern __shared__ unsigned char smem_raw[];
  Cplx<double>* work = reinterpret_cast<Cplx<double>*>(smem_raw);
  const int i = blockIdx.x, tid = threadIdx.x, nth = blockDim.x;
  const int M = nz + 1;
  const double* row = src + (int64_t)i * nz;
  any_dft_stage(work, a, tid, nth);
  for (int j = tid; j < nz; j += nth) {
    const double x = row[j];
    packed_y(work, j + 1) = x;
    packed_y(work, 2 * M - 1 - j) = -x;
  }
  if (tid == 0) {
    packed_y(work, 0) = 0.0;
    packed_y(work, M) = 0.0;
  }
  __syncthreads();
  any_dft<double, PER>(work, a, -1, tid, nth);
  double* out = dst + (int64_t)i * nz;
  for (int k = tid + 1; k <= nz; k += nth) {
    const Cplx<double> zk = work[k], zc = cconj(work[M - k]);
    const Cplx<double> s = cadd(zk, zc), d = csub(zk, zc);
    const Cplx<double> ed = cmul(em[k], d);
    // -Im(Y_k) / 2 with Y_k = s / 2 - i ed / 2
    out[k - 1] = (ed.x - s.y) * norm;
  }
}

// Every kernel above may take more than the default 64 KiB of dynamic LDS
// (a length-8192 FP64 row is 128 KiB of the CU's 160 KiB): each is allowed
// the device's opt-in maximum per workgroup less its own static LDS.  A
// refused request leaves the default limit (a launch above it then fails
// with a launch error); the error it records is cleared here so that the
// next launch check does not report it.
static void allow_large_lds() {
  static std::once_flag once;
  std::call_once(once, [] {
    int dev = 0, optin = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) !=
            hipSuccess ||
        optin <= 0) {
      (void)hipGetLastError();
      return;
    }
    const void* ks[] = {
        (const void*)any_bhat_kernel<1>, (const void*)any_bhat_kernel<2>,
        (const void*)any_bhat_kernel<4>,
        (const void*)any_dct2_rows_kernel<double, 1, false>,
        (const void*)any_dct2_rows_kernel<double, 2, false>,
        (const void*)any_dct2_rows_kernel<double, 4, false>,
        (const void*)any_dct2_rows_kernel<double, 1, true>,
        (const void*)any_dct2_rows_kernel<double, 2, true>,
        (const void*)any_dct2_rows_kernel<double, 4, true>,
        (const void*)any_dct2_rows_kernel<float, 1, false>,
        (const void*)any_dct2_rows_kernel<float, 2, false>,
        (const void*)any_dct2_rows_kernel<float, 4, false>,
        (const void*)any_dct2_rows_kernel<float, 1, true>,
        (const void*)any_dct2_rows_kernel<float, 2, true>,
        (const void*)any_dct2_rows_kernel<float, 4, true>,
        (const void*)any_dct3_rows_kernel<double, 1>, (const void*)any_dct3_rows_kernel<double, 2>,
        (const void*)any_dct3_rows_kernel<double, 4>, (const void*)any_dct3_rows_kernel<float, 1>,
        (const void*)any_dct3_rows_kernel<float, 2>, (const void*)any_dct3_rows_kernel<float, 4>,
        (const void*)any_rfft_rows_kernel<1>, (const void*)any_rfft_rows_kernel<2>,
        (const void*)any_rfft_rows_kernel<4>, (const void*)any_irfft_rows_kernel<1>,
        (const void*)any_irfft_rows_kernel<2>, (const void*)any_irfft_rows_kernel<4>,
        (const void*)any_dst1_rows_kernel<1>, (const void*)any_dst1_rows_kernel<2>,
        (const void*)any_dst1_rows_kernel<4>,
        (const void*)soft_batch_dct2_kernel<double, 1>,
        (const void*)soft_batch_dct2_kernel<double, 2>,
        (const void*)soft_batch_dct2_kernel<double, 4>,
        (const void*)soft_batch_dct3_kernel<double, 1>,
        (const void*)soft_batch_dct3_kernel<double, 2>,
        (const void*)soft_batch_dct3_kernel<double, 4>,
        (const void*)soft_batch_dct2_kernel<float, 1>,
        (const void*)soft_batch_dct2_kernel<float, 2>,
        (const void*)soft_batch_dct2_kernel<float, 4>,
        (const void*)soft_batch_dct3_kernel<float, 1>,
        (const void*)soft_batch_dct3_kernel<float, 2>,
        (const void*)soft_batch_dct3_kernel<float, 4>};
    int limit = optin;
    for (const void* k : ks) {
      hipFuncAttributes fa{};
      int allowed = 65536;
      if (hipFuncGetAttributes(&fa, k) == hipSuccess) {
        const int want = optin - (int)fa.sharedSizeBytes;
        if (want > 65536 &&
            hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, want) == hipSuccess)
          allowed = want;
      }
      (void)hipGetLastError();
      limit = std::min(limit, allowed);
    }
    g_any_lds_limit = limit;
  });
}

// ------------------------------------------------------------------ launchers

#define AXS_ANY_PER(PERV, CALL) \
  do {                       \
    if ((PERV) == 1) {       \
      constexpr int P = 1;   \
      CALL;                  \
    } else if ((PERV) == 2) { \
      constexpr int P = 2;   \
      CALL;                  \
    } else {                 \
      constexpr int P = 4;   \
      CALL;                  \
    }                        \
  } while (0)

template <typename T>
hipError_t launch_any_dct2_rows(const StokesPlan& pl, const T* src, T* tapered, T* dst,
                                const double* rscale, const double* wz, const double* wr, int w,
                                bool taper, int batch, const StepFinalize* fin, hipStream_t s) {
  const AnyDft& D = *pl.any;
  const AnyDctTables tb{D.args, pl.ge2, pl.e1, pl.e2, pl.any_half ? 1 : 0};
  const StepFinalize f = fin ? *fin : StepFinalize{};
  const dim3 grid(pl.nr + (f.partials ? 1 : 0), batch);
  const size_t shm = D.lds_bytes<T>();
  const int nth = D.nth;
  if (taper)
    AXS_ANY_PER(D.per, (any_dct2_rows_kernel<T, P, true><<<grid, nth, shm, s>>>(
                       src, tapered, dst, rscale, wz, wr, pl.nr, pl.nz, w, tb, f)));
  else
    AXS_ANY_PER(D.per, (any_dct2_rows_kernel<T, P, false><<<grid, nth, shm, s>>>(
                       src, nullptr, dst, rscale, nullptr, nullptr, pl.nr, pl.nz, 0, tb, f)));
  return hipGetLastError();
}

template <typename T>
hipError_t launch_any_dct3_rows(const StokesPlan& pl, const T* src, T* dst, int rows, int batch,
                                hipStream_t s) {
  const AnyDft& D = *pl.any;
  const AnyDctTables tb{D.args, pl.ge2, pl.e1, pl.e2, pl.any_half ? 1 : 0};
  AXS_ANY_PER(D.per, (any_dct3_rows_kernel<T, P><<<dim3(rows, batch), D.nth, D.lds_bytes<T>(), s>>>(
                     src, dst, rows, pl.nz, tb)));
  return hipGetLastError();
}

hipError_t launch_any_rfft_rows(const StokesPlan& pl, const double* src, double* dst,
                                const double* rscale, int rows, bool inverse, hipStream_t s) {
  const AnyDft& D = *pl.any;
  const size_t shm = D.lds_bytes<double>();
  if (inverse)
    AXS_ANY_PER(D.per, (any_irfft_rows_kernel<P><<<rows, D.nth, shm, s>>>(src, dst, pl.nz, D.args)));
  else
    AXS_ANY_PER(D.per, (any_rfft_rows_kernel<P><<<rows, D.nth, shm, s>>>(src, dst, rscale, pl.nz,
                                                                      D.args)));
  return hipGetLastError();
}

hipError_t launch_any_dst1_rows(const AnyDft& D, const Cplx<double>* em, const double* src,
                                double* dst, int rows, int nz, hipStream_t s) {
  const double norm = std::sqrt(2.0 / (double)(nz + 1)) / 4.0;
  AXS_ANY_PER(D.per, (any_dst1_rows_kernel<P><<<rows, D.nth, D.lds_bytes<double>(), s>>>(
                     src, dst, nz, D.args, em, norm)));
  return hipGetLastError();
}

template <typename B>
hipError_t launch_soft_batch_dct2(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                  hipStream_t s) {
  const dim3 grid(d.max_nr, d.ncase);
  AXS_ANY_PER(d.any_per,
              (soft_batch_dct2_kernel<B, P><<<grid, d.any_nth, d.any_lds, s>>>(cases)));
  return hipGetLastError();
}

template <typename B>
hipError_t launch_soft_batch_dct3(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                  hipStream_t s) {
  const dim3 grid(d.max_nr, d.ncase);
  AXS_ANY_PER(d.any_per,
              (soft_batch_dct3_kernel<B, P><<<grid, d.any_nth, d.any_lds, s>>>(cases)));
  return hipGetLastError();
}

template hipError_t launch_soft_batch_dct2<double>(const SoftCaseT<double>*, const SoftBatchDims&,
                                                   hipStream_t);
template hipError_t launch_soft_batch_dct2<float>(const SoftCaseT<float>*, const SoftBatchDims&,
                                                  hipStream_t);
template hipError_t launch_soft_batch_dct3<double>(const SoftCaseT<double>*, const SoftBatchDims&,
                                                   hipStream_t);
template hipError_t launch_soft_batch_dct3<float>(const SoftCaseT<float>*, const SoftBatchDims&,
                                                  hipStream_t);

#undef AXS_ANY_PER

template hipError_t launch_any_dct2_rows<double>(const StokesPlan&, const double*, double*,
                                                 double*, const double*, const double*,
                                                 const double*, int, bool, int,
                                                 const StepFinalize*, hipStream_t);
template hipError_t launch_any_dct2_rows<float>(const StokesPlan&, const float*, float*, float*,
                                                const double*, const double*, const double*, int,
                                                bool, int, const StepFinalize*, hipStream_t);
template hipError_t launch_any_dct3_rows<double>(const StokesPlan&, const double*, double*, int,
                                                 int, hipStream_t);
template hipError_t launch_any_dct3_rows<float>(const StokesPlan&, const float*, float*, int,
                                                int, hipStream_t);

}  // namespace axs
