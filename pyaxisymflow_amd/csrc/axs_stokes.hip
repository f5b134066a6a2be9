// Stream-function Poisson solve of FastDiagonalisationStokesSolver
// (kernels/FastDiagonalisationStokesSolver.py:6-156), restated as
//   1. row pass:    what = DCT-II_z(r * rhs)          (one workgroup per row)
//   2. column pass: (A_r + lambda_k I) psihat_k = what_k for every z-mode k
//   3. row pass:    psi = DCT-II^{-1}_z(psihat)
// A_r = P_r - D_r is the reference's radial matrix (tridiagonal, built with
// the same floating-point operations as :41-86); P_z (Neumann) has the
// DCT-II eigenvectors cos(pi k (2n+1)/(2 nz)) and eigenvalues
// lambda_k = (4/dx^2) sin^2(pi k / (2 nz)).
// The transforms are a different algorithm from the reference's dense
// route, so FMA contraction is allowed here (the library default is off).
#pragma clang fp contract(fast)

#include "axs_fft1024.hpp"
#include "axs_fused.hpp"
#include "axs_softbatch.hpp"
#include "axs_stokes.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <sys/stat.h>
#include <tuple>
#include <vector>

namespace axs {

// ------------------------------------------------------------------ host plan

static bool factor_radices(int M, FftPlanArgs* p) {
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

template <typename V>
static hipError_t upload(V** dev, const std::vector<V>& host) {
  hipError_t e = hipMalloc(reinterpret_cast<void**>(dev), sizeof(V) * std::max<size_t>(host.size(), 1));
  if (e != hipSuccess) return e;
  return hipMemcpy(*dev, host.data(), sizeof(V) * host.size(), hipMemcpyHostToDevice);
}

StokesPlan::~StokesPlan() {
  for (void* p : {(void*)tw, (void*)e1, (void*)e2, (void*)a, (void*)b, (void*)c,
                  (void*)lam, (void*)rvec, (void*)cseg, (void*)spec, (void*)ge2})
    if (p) (void)hipFree(p);
  delete any;
}

bool stokes_nz_supported(int nz) {
  if (nz < 2) return false;
  FftPlanArgs fp;
  if (nz % 2 == 0 && nz <= 8192 && factor_radices(nz / 2, &fp)) return true;
  bool blue = false;
  return any_dft_length(nz, &blue) <= kAnyMaxL;
}

int StokesPlan::tri_threads() const { return tri_K * tri_P; }

// c'_{pL-1} of the forward elimination at every segment start p L, per mode
// (the pivot recurrence depends only on the matrix, not on the data).
std::vector<double> pivot_table(const std::vector<double>& a, const std::vector<double>& b,
                                       const std::vector<double>& c,
                                       const std::vector<double>& lam, int nr, int nz, int L) {
  const int P = (nr + L - 1) / L;
  std::vector<double> t((size_t)P * nz, 0.0);
  for (int k = 0; k < nz; ++k) {
    double cp = 0.0;
    for (int i = 0; i < nr; ++i) {
      if (i % L == 0) t[(size_t)(i / L) * nz + k] = cp;
      const double m = (b[i] + lam[k]) - a[i] * cp;
      cp = c[i] / m;
    }
  }
  return t;
}

// Autotuner of the column pass: times every admissible (modes per
// workgroup K, rows per segment L) on this device with the plan's own
// matrices and keeps the fastest; results are cached per (nr, nz).
// On-disk record of the column-pass layouts the timed search has chosen, so
// that later processes on the same kind of device take the same layout (the
// radial solve's round-off depends on it) without searching again.  One line
// per choice: "<arch> <nr> <nz> <key> <L> <K> <G>"; the first line for a
// grid wins.  AXS_TUNE_CACHE names the file ("off": none); by default
// $XDG_CACHE_HOME/axisym-hip/tridiag.txt or ~/.cache/axisym-hip/tridiag.txt.
static std::string tune_cache_path() {
  if (const char* p = std::getenv("AXS_TUNE_CACHE")) return std::strcmp(p, "off") ? p : "";
  std::string dir;
  if (const char* x = std::getenv("XDG_CACHE_HOME"))
    dir = x;
  else if (const char* h = std::getenv("HOME"))
    dir = std::string(h) + "/.cache";
  else
    return "";
  dir += "/axisym-hip";
  (void)::mkdir(dir.substr(0, dir.rfind('/')).c_str(), 0755);
  (void)::mkdir(dir.c_str(), 0755);
  return dir + "/tridiag.txt";
}

static std::string device_arch() {
  int dev = 0;
  hipDeviceProp_t prop{};
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    (void)hipGetLastError();
    return "unknown";
  }
  std::string a = prop.gcnArchName;
  for (char& c : a)
    if (c == ' ') c = '_';
  return a + "_cu" + std::to_string(prop.multiProcessorCount);
}

static bool tune_cache_read(const std::string& path, const std::string& arch, int nr, int nz,
                            int key, int* L, int* K, int* G) {
  if (path.empty()) return false;
  std::FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  char a[128];
  int r, z, k, l, kk, g;
  bool found = false;
  while (!found && std::fscanf(f, "%127s %d %d %d %d %d %d", a, &r, &z, &k, &l, &kk, &g) == 7)
    if (arch == a && r == nr && z == nz && k == key) {
      *L = l;
      *K = kk;
      *G = g;
      found = true;
    }
  std::fclose(f);
  return found;
}

static void tune_cache_write(const std::string& path, const std::string& arch, int nr, int nz,
                             int key, int L, int K, int G) {
  if (path.empty()) return;
  if (std::FILE* f = std::fopen(path.c_str(), "a")) {
    std::fprintf(f, "%s %d %d %d %d %d %d\n", arch.c_str(), nr, nz, key, L, K, G);
    std::fclose(f);
  }
}

hipError_t StokesPlan::tune_tridiag(const std::vector<double>& ha, const std::vector<double>& hb,
                                    const std::vector<double>& hc,
                                    const std::vector<double>& hlam) {
  static std::mutex mu;
  struct Choice {
    int K, L;
    bool G;
  };
  static std::map<std::tuple<int, int, int>, Choice> cache;
  Choice best{tri_K, tri_L, tri_G};
  bool cached = false;
  // a layout is usable when its workgroup and LDS fit
  auto usable = [&](int L, int K, int G) {
    if ((L != 4 && L != 8 && L != 16) || K < 1 || G < 0 || G > 1) return false;
    const int P = (nr + L - 1) / L;
    return K * P <= 1024 && tri_lds_bytes(nr, L, K, P, G == 1) <= 65536;
  };
  // AXS_TRIDIAG_LK="L,K" pins the choice (the search times the candidates,
  // so two processes may pick different ones when timings tie, and the
  // radial solve's round-off depends on the choice)
  if (const char* pin = std::getenv("AXS_TRIDIAG_LK")) {  // "L,K" or "L,K,G" (G = 1: global)
    int L = 0, K = 0, G = 0;
    const int got = std::sscanf(pin, "%d,%d,%d", &L, &K, &G);
    if (got < 2 || !usable(L, K, G)) return hipErrorInvalidValue;
    best = {K, L, G == 1};
    cached = true;
  }
  const std::string path = cached ? std::string() : tune_cache_path();
  const std::string arch = cached ? std::string() : device_arch();
  if (!cached) {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find({nr, nz, tune_key});
    if (it != cache.end()) {
      best = it->second;
      cached = true;
    } else {
      int L = 0, K = 0, G = 0;
      if (tune_cache_read(path, arch, nr, nz, tune_key, &L, &K, &G) && usable(L, K, G)) {
        best = {K, L, G == 1};
        cache[{nr, nz, tune_key}] = best;
        cached = true;
      }
    }
  }
  if (!cached) {
    // Each candidate is timed right after a forward row pass that writes the
    // spectrum, as in the step: timed alone, back to back, the radial pass
    // finds its modes still in each XCD's L2 and ranks the layouts
    // differently (a layout 10% slower in the step measured fastest alone).
    // A plan without row tables (the implicit diffusion's radial plan, whose
    // rows are DST-I passes of its own) times the radial pass alone.
    const bool row_tables = general ? any != nullptr : (tw && e1 && e2);
    const size_t cells = (size_t)nr * nz;
    double* rows_in = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&rows_in), sizeof(double) * cells);
    if (e != hipSuccess) return e;
    if ((e = hipMemset(rows_in, 0, sizeof(double) * cells)) != hipSuccess) {
      (void)hipFree(rows_in);
      return e;
    }
    if ((e = hipMemset(spec, 0, sizeof(double) * cells)) != hipSuccess) {
      (void)hipFree(rows_in);
      return e;
    }
    auto in_step = [&]() {
      hipError_t le =
          row_tables ? launch_dct2_rows<double>(*this, rows_in, spec, nullptr, nr, 0) : hipSuccess;
      return le == hipSuccess ? launch_tridiag<double>(*this, spec, 0) : le;
    };
    hipEvent_t t0, t1;
    if ((e = hipEventCreate(&t0)) != hipSuccess || (e = hipEventCreate(&t1)) != hipSuccess) {
      (void)hipFree(rows_in);
      return e;
    }
    float best_ms = 1e30f;
    double* saved = cseg;
    const int saved_k = tri_K, saved_l = tri_L, saved_p = tri_P;
    const bool saved_g = tri_G;
    for (int L : {4, 8, 16}) {
      const int P = (nr + L - 1) / L;
      if (P > 512) continue;
      double* tab = nullptr;
      if ((e = upload(&tab, pivot_table(ha, hb, hc, hlam, nr, nz, L))) != hipSuccess) break;
      for (int K = 1; K <= 64; K *= 2)
      for (int G = 0; G < 2; ++G) {
        if (K * P > 1024 || K * P < 64) continue;
        if (tri_lds_bytes(nr, L, K, P, G == 1) > 65536) continue;
        cseg = tab;
        tri_K = K;
        tri_L = L;
        tri_P = P;
        tri_G = G == 1;
        // the fastest of three batches of five (row pass, radial pass)
        // pairs (one batch is noisy enough to rank neighbours wrongly)
        float ms = 1e30f;
        for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) e = in_step();
        for (int trial = 0; trial < 3 && e == hipSuccess; ++trial) {
          float t = 0.f;
          e = hipEventRecord(t0, 0);
          for (int rep = 0; rep < 5 && e == hipSuccess; ++rep) e = in_step();
          if (e == hipSuccess) e = hipEventRecord(t1, 0);
          if (e == hipSuccess) e = hipEventSynchronize(t1);
          if (e == hipSuccess) e = hipEventElapsedTime(&t, t0, t1);
          if (e == hipSuccess && t < ms) ms = t;
        }
        if (e != hipSuccess) break;
        if (ms < best_ms) {
          best_ms = ms;
          best = {K, L, G == 1};
        }
      }
      (void)hipFree(tab);
      if (e != hipSuccess) break;
    }
    cseg = saved;
    tri_K = saved_k;
    tri_L = saved_l;
    tri_P = saved_p;
    tri_G = saved_g;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipFree(rows_in);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> g(mu);
    // another process may have recorded this grid meanwhile: its choice wins
    int L = 0, K = 0, G = 0;
    if (tune_cache_read(path, arch, nr, nz, tune_key, &L, &K, &G) && usable(L, K, G))
      best = {K, L, G == 1};
    else
      tune_cache_write(path, arch, nr, nz, tune_key, best.L, best.K, best.G ? 1 : 0);
    cache[{nr, nz, tune_key}] = best;
  }
  tri_G = best.G;
  if (best.K == tri_K && best.L == tri_L) return hipSuccess;
  double* tab = nullptr;
  hipError_t e = upload(&tab, pivot_table(ha, hb, hc, hlam, nr, nz, best.L));
  if (e != hipSuccess) return e;
  (void)hipFree(cseg);
  cseg = tab;
  tri_K = best.K;
  tri_L = best.L;
  tri_P = (nr + tri_L - 1) / tri_L;
  return hipSuccess;
}

hipError_t StokesPlan::init(int nr_, int nz_, double dx_, int bc_) {
  bc = bc_;
  hipError_t e = init_transforms(nr_, nz_, dx_);
  if (e != hipSuccess) return e;
  const std::vector<double>& h_r = host_r;
  // A_r = P_r - D_r (:45-65, Neumann BCs :79-86), entries as the reference forms them.
  const double inv_dx2 = 1.0 / dx / dx;
  const double inv_2dx = 1.0 / 2.0 / dx;
  std::vector<double> h_a(nr, 0.0), h_b(nr), h_c(nr, 0.0);
  for (int i = 0; i < nr; ++i) {
    const double P_sub = inv_dx2 * -1.0, P_sup = inv_dx2 * -1.0;
    const double D_sub = (inv_2dx * 1.0) / h_r[i], D_sup = (inv_2dx * -1.0) / h_r[i];
    h_b[i] = inv_dx2 * 2.0;
    if (i > 0) h_a[i] = P_sub - D_sub;
    if (i < nr - 1) h_c[i] = P_sup - D_sup;
  }
  if (bc != 2) {                           // Neumann at r_max (:79-86)
    h_b[nr - 1] = inv_dx2;                 // P_r[-1,-1] = inv_dx2
    if (nr > 1) h_a[nr - 1] = inv_dx2 * -1.0 - 0.0;  // D_r[-1,-2] = 0
  }
  std::vector<double> h_lam(nz);
  if (bc == 0) {  // DCT-II modes of the Neumann P_z
    for (int k = 0; k < nz; ++k) {
      const double s = std::sin(M_PI * (double)k / (2.0 * nz));
      h_lam[k] = (4.0 / (dx * dx)) * s * s;
    }
  } else {  // DFT modes of the periodic P_z in the packed real layout
    for (int p = 0; p < nz; ++p) {
      // packed layout [X_0, Re X_1, Im X_1, ..., (X_{nz/2} when nz is even)]
      const int k = (p == 0) ? 0 : (p == nz - 1 && nz % 2 == 0) ? nz / 2 : (p + 1) / 2;
      const double s = std::sin(M_PI * (double)k / (double)nz);
      h_lam[p] = (4.0 / (dx * dx)) * s * s;
    }
  }
  return set_tridiag(h_a, h_b, h_c, h_lam);
}

hipError_t StokesPlan::init_transforms(int nr_, int nz_, double dx_) {
  nr = nr_;
  nz = nz_;
  dx = dx_;
  general = force_general || !(nz % 2 == 0 && factor_radices(nz / 2, &fft));
  hipError_t e;
  if (general) {
    // the cosine rows of an even nz take the half-length DFT; the periodic
    // real-DFT rows take length nz
    any_half = bc == 0 && nz % 2 == 0;
    any = new (std::nothrow) AnyDft();
    if (!any) return hipErrorOutOfMemory;
    if ((e = any->init(any_half ? nz / 2 : nz)) != hipSuccess) return e;
  }
  if (general && !any_half) {
    std::vector<Cplx<double>> h_ge2(nz);
    for (int k = 0; k < nz; ++k) {
      const double ang = -M_PI * (double)k / (2.0 * nz);
      h_ge2[k] = {std::cos(ang), std::sin(ang)};
    }
    if ((e = upload(&ge2, h_ge2)) != hipSuccess) return e;
  }
  const int M = nz / 2;
  // FFT twiddles e^{-2 pi i t/M}
  std::vector<Cplx<double>> h_tw(M), h_e1(M + 1), h_e2(M + 1);
  for (int t = 0; t < M; ++t) {
    const double ang = -2.0 * M_PI * (double)t / (double)M;
    h_tw[t] = {std::cos(ang), std::sin(ang)};
  }
  for (int k = 0; k <= M; ++k) {
    const double a1 = -2.0 * M_PI * (double)k / (double)nz;
    const double a2 = -M_PI * (double)k / (2.0 * nz);
    h_e1[k] = {std::cos(a1), std::sin(a1)};
    h_e2[k] = {std::cos(a2), std::sin(a2)};
  }
  // Radial coordinate, FastDiagonalisationStokesSolver.py:20-24:
  // numpy.linspace(dx/2, nr*dx - dx/2, nr) = start + i*step, last = stop.
  std::vector<double> h_r(nr);
  const double start = dx / 2, stop = nr * dx - dx / 2;
  const double step = (nr > 1) ? (stop - start) / (nr - 1) : 0.0;
  for (int i = 0; i < nr; ++i) h_r[i] = i * step + start;
  if (nr > 1) h_r[nr - 1] = stop;
  host_r = h_r;
  if ((e = upload(&tw, h_tw)) != hipSuccess) return e;
  if ((e = upload(&e1, h_e1)) != hipSuccess) return e;
  if ((e = upload(&e2, h_e2)) != hipSuccess) return e;
  return upload(&rvec, h_r);
}

hipError_t StokesPlan::set_tridiag(const std::vector<double>& h_a, const std::vector<double>& h_b,
                                   const std::vector<double>& h_c,
                                   const std::vector<double>& h_lam) {
  // Segment layout of the column pass and c'_{s-1} at every segment start.
  // Segments of L = 8 rows (16 for very tall grids) keep the per-thread
  // registers small; K modes per workgroup so that K * P <= 1024 and, when
  // the grid allows, at least 512 workgroups are in flight.
  tri_L = 8;
  tri_P = (nr + tri_L - 1) / tri_L;
  while (tri_P > 256) {
    tri_L *= 2;
    tri_P = (nr + tri_L - 1) / tri_L;
  }
  tri_K = 1;
  while (tri_K * 2 * tri_P <= 1024 && tri_K < 64 && nz / (tri_K * 2) >= 512) tri_K *= 2;
  if (tri_K < 2 && 2 * tri_P <= 1024) tri_K = 2;
  // keep the column kernel's LDS (scan buffers + coefficients) within 64 KiB
  if (force_L > 0) {
    tri_L = force_L;
    tri_P = (nr + tri_L - 1) / tri_L;
    if (tri_P > 1024) return hipErrorInvalidValue;
    tri_K = std::max(1, 1024 / tri_P);
  }
  tri_G = tri_lds_bytes(nr, tri_L, tri_K, tri_P, false) > 65536;
  while (tri_K > 1 && tri_lds_bytes(nr, tri_L, tri_K, tri_P, tri_G) > 65536) tri_K /= 2;
  if (tri_lds_bytes(nr, tri_L, tri_K, tri_P, tri_G) > 65536) return hipErrorInvalidValue;
  std::vector<double> h_cseg = pivot_table(h_a, h_b, h_c, h_lam, nr, nz, tri_L);
  hipError_t e;
  if ((e = upload(&a, h_a)) != hipSuccess) return e;
  if ((e = upload(&b, h_b)) != hipSuccess) return e;
  if ((e = upload(&c, h_c)) != hipSuccess) return e;
  if ((e = upload(&lam, h_lam)) != hipSuccess) return e;
  if ((e = upload(&cseg, h_cseg)) != hipSuccess) return e;
  e = hipMalloc(reinterpret_cast<void**>(&spec), sizeof(double) * (size_t)nr * nz);
  if (e != hipSuccess) return e;
  if (force_L > 0) return hipSuccess;
  if (std::getenv("AXS_TRIDIAG_LK") ||
      ((size_t)nr * nz >= (size_t(1) << 18) && !std::getenv("AXS_NO_TUNE")))
    return tune_tridiag(h_a, h_b, h_c, h_lam);
  return hipSuccess;
}

// ------------------------------------------------------------------ kernels

// Periodic z (bc_type *_periodic_along_z): real DFT of each row through the
// complex FFT of its M = nz / 2 packed pairs z_m = x_{2m} + i x_{2m+1}:
//   X_k = (Z_k + conj Z_{M-k}) / 2 - i W_N^k (Z_k - conj Z_{M-k}) / 2,  k = 0..M,
// stored as the packed real row [X_0, Re X_1, Im X_1, ..., Re X_{M-1},
// Im X_{M-1}, X_M] (X_0 and X_M are real).  Every entry is one real
// unknown of its mode's radial system.  W_N^k = e1[k].
template <int PER>
__global__ void __launch_bounds__(256) rfft_rows_kernel(
    const double* __restrict__ src, double* __restrict__ dst, const double* __restrict__ rscale,
    int nz, FftPlanArgs fp, const Cplx<double>* __restrict__ tw,
    const Cplx<double>* __restrict__ e1) {
  extern __shared__ unsigned char smem_raw[];
  const int M = fp.M;
  Cplx<double>* buf = reinterpret_cast<Cplx<double>*>(smem_raw);
  const int i = blockIdx.x;
  const double* row = src + (int64_t)i * nz;
  const double sc = rscale ? rscale[i] : 1.0;
  for (int j = threadIdx.x; j < nz; j += blockDim.x) packed_component(buf, j) = sc * row[j];
  __syncthreads();
  fft_lds<double, PER>(buf, fp, tw, -1, threadIdx.x, blockDim.x);
  double* out = dst + (int64_t)i * nz;
  for (int k = threadIdx.x; k <= M; k += blockDim.x) {
    const Cplx<double> zk = buf[k == M ? 0 : k], zm = buf[k == 0 ? 0 : M - k];
    const Cplx<double> a = cadd(zk, cconj(zm)), b = csub(zk, cconj(zm));
    const Cplx<double> c = cmul(e1[k], b);
    const double xr = 0.5 * (a.x + c.y), xi = 0.5 * (a.y - c.x);
    if (k == 0)
      out[0] = xr;
    else if (k == M)
      out[nz - 1] = xr;
    else {
      out[2 * k - 1] = xr;
      out[2 * k] = xi;
    }
  }
}

// Inverse: E_k = (X_k + conj X_{M-k}) / 2, O_k = conj(W_N^k) (X_k - conj X_{M-k}) / 2,
// Z_k = E_k + i O_k, inverse FFT, x_{2m} + i x_{2m+1} = z_m / M.
template <int PER>
__global__ void __launch_bounds__(256) irfft_rows_kernel(
    const double* __restrict__ src, double* __restrict__ dst, int nz, FftPlanArgs fp,
    const Cplx<double>* __restrict__ tw, const Cplx<double>* __restrict__ e1) {
  extern __shared__ unsigned char smem_raw[];
  const int M = fp.M;
  Cplx<double>* buf = reinterpret_cast<Cplx<double>*>(smem_raw);
  const int i = blockIdx.x;
  const double* row = src + (int64_t)i * nz;
  auto X = [&](int k) -> Cplx<double> {
    if (k == 0) return {row[0], 0.0};
    if (k == M) return {row[nz - 1], 0.0};
    return {row[2 * k - 1], row[2 * k]};
  };
  for (int k = threadIdx.x; k < M; k += blockDim.x) {
    const Cplx<double> xk = X(k), xm = X(M - k);
    const Cplx<double> e = {0.5 * (xk.x + xm.x), 0.5 * (xk.y - xm.y)};
    const Cplx<double> d = csub(xk, cconj(xm));
    const Cplx<double> o = cmul(cconj(e1[k]), Cplx<double>{0.5 * d.x, 0.5 * d.y});
    buf[k] = {e.x - o.y, e.y + o.x};
  }
  __syncthreads();
  fft_lds<double, PER>(buf, fp, tw, +1, threadIdx.x, blockDim.x);
  double* out = dst + (int64_t)i * nz;
  const double invM = 1.0 / (double)M;
  for (int j = threadIdx.x; j < nz; j += blockDim.x) out[j] = packed_component(buf, j) * invM;
}

// Row pass: x_j = scale_i * src[i, j]  ->  X = DCT-II(x)  ->  dst[i, :].
// The row is read coalesced and scattered into the packed LDS sequence; the
// post-processing twiddles are fetched before the FFT starts.
template <typename T, int PER>
__global__ void __launch_bounds__(256) dct2_rows_kernel(
    const T* __restrict__ src, T* __restrict__ dst, const double* __restrict__ rscale,
    int nz, FftPlanArgs fp, const Cplx<double>* __restrict__ tw,
    const Cplx<double>* __restrict__ e1, const Cplx<double>* __restrict__ e2) {
  extern __shared__ unsigned char smem_raw[];
  const int M = fp.M;
  Cplx<T>* buf = reinterpret_cast<Cplx<T>*>(smem_raw);
  const int i = blockIdx.x;
  const T* row = src + (int64_t)i * nz;
  const T s = rscale ? T(rscale[i]) : T(1);
  constexpr int NJ = 4 * PER;  // row entries per thread (4 * PER * blockDim >= nz)
  T v[NJ];
#pragma unroll
  for (int q = 0; q < NJ; ++q) {
    const int j = threadIdx.x + q * blockDim.x;
    if (j < nz) v[q] = row[j];
  }
  PostTwiddles<PER + 1> pt;
  dct2_post_twiddles<PER + 1>(pt, M, e1, e2, threadIdx.x, blockDim.x);
#pragma unroll
  for (int q = 0; q < NJ; ++q) {
    const int j = threadIdx.x + q * blockDim.x;
    if (j < nz) packed_component(buf, dct_dst_index(j, nz)) = s * v[q];
  }
  __syncthreads();
  fft_lds<T, PER>(buf, fp, tw, -1, threadIdx.x, blockDim.x);
  T* out = dst + (int64_t)i * nz;
  dct2_post<T, PER + 1>(buf, M, pt, [&](int k, T val) { out[k] = val; }, threadIdx.x, blockDim.x);
}

// Fused first pass of a step: the reference's kill taper
// (kill_boundary_vorticity_sine.py:4-27: z ends of every row, then rows
// nr-w..nr-1 from row nr-w, then row 0 = 0) applied on the fly to the
// untouched field `src`; the tapered row goes to `tapered` (when not null)
// and r_i * row goes through the DCT-II into `dst`.  Reading
// src and writing a separate buffer keeps rows independent (the top rows
// read row nr-w).  wz: 2w weights (left block, right block), wr: w weights.
// One row of the tapered forward pass (src, tapered, dst at the case).
// Threads tid of nth (all of them call it: the FFT has block barriers)
// work in buf, M complex entries of LDS.  STAGE: the FFT reads its twiddles
// from tl (M entries of LDS), filled from tw by threads wtid of wnth (every
// row of a workgroup passes the same tl and its workgroup-wide thread
// index); else from tw in global memory (tl, wtid, wnth unused).
template <typename T, int PER, bool STAGE>
__device__ __forceinline__ void taper_dct2_row(
    const T* __restrict__ src, T* __restrict__ tapered, T* __restrict__ dst,
    const double* __restrict__ rscale, const double* __restrict__ wz,
    const double* __restrict__ wr, int nr, int nz, int w, const FftPlanArgs& fp,
    const Cplx<double>* __restrict__ tw, const Cplx<double>* __restrict__ e1,
    const Cplx<double>* __restrict__ e2, int i, Cplx<T>* buf, Cplx<T>* tl, int tid, int nth,
    int wtid, int wnth) {
  const int M = fp.M;
  const int srow = (i >= nr - w) ? nr - w : i;
  const T* row = src + (int64_t)srow * nz;
  constexpr int NJ = 4 * PER;
  T v[NJ];
#pragma unroll
  for (int q = 0; q < NJ; ++q) {
    const int j = tid + q * nth;
    if (j < nz) v[q] = row[j];
  }
  const T left = row[w - 1], right = row[nz - w];
  PostTwiddles<PER + 1> pt;
  dct2_post_twiddles<PER + 1>(pt, M, e1, e2, tid, nth);
  if constexpr (STAGE) stage_twiddles(tl, tw, M, wtid, wnth);
  const T rs = T(rscale[i]);
  const T rowf = (i == 0) ? T(0) : ((i >= nr - w) ? T(wr[i - (nr - w)]) : T(1));
  T* out_t = tapered ? tapered + (int64_t)i * nz : nullptr;
#pragma unroll
  for (int q = 0; q < NJ; ++q) {
    const int j = tid + q * nth;
    if (j < nz) {
      T t = v[q];
      if (j < w)
        t = T(wz[j]) * left;
      else if (j >= nz - w)
        t = T(wz[w + (j - (nz - w))]) * right;
      if (i >= nr - w) t = rowf * t;
      if (i == 0) t = T(0);
      if (out_t) out_t[j] = t;
      packed_component(buf, dct_dst_index(j, nz)) = rs * t;
    }
  }
  __syncthreads();
  if constexpr (STAGE)
    fft_lds<T, PER>(buf, fp, static_cast<const Cplx<T>*>(tl), -1, tid, nth);
  else
    fft_lds<T, PER>(buf, fp, tw, -1, tid, nth);
  T* out = dst + (int64_t)i * nz;
  dct2_post<T, PER + 1>(buf, M, pt, [&](int k, T val) { out[k] = val; }, tid, nth);
}

// Row pass: x = DCT-II^{-1}(src[i, :]) -> dst[i, :], stored coalesced by
// gathering each column's value from the packed LDS sequence.

template <typename T, int PER>
__global__ void __launch_bounds__(256) taper_dct2_rows_kernel(
    const T* __restrict__ src, T* __restrict__ tapered, T* __restrict__ dst,
    const double* __restrict__ rscale, const double* __restrict__ wz,
    const double* __restrict__ wr, int nr, int nz, int w, FftPlanArgs fp,
    const Cplx<double>* __restrict__ tw, const Cplx<double>* __restrict__ e1,
    const Cplx<double>* __restrict__ e2, StepFinalize fin) {
  if (fin.partials && blockIdx.x == nr) {  // completes the previous step
    __shared__ double fin_red[256];
    const int cb = blockIdx.y;
    const RigidConfig& cfg = fin.cfgs[cb];
    finalize_pending(fin.partials + (int64_t)cb * fin.count, fin.count,
                     fin.slots + (int64_t)cb * kSlotsPerCase, fin.sc + (int64_t)cb * AXS_RIGID_SCALARS,
                     fin.drag ? fin.drag + (int64_t)cb * cfg.history : nullptr, cfg, fin_red);
    return;
  }
  const int64_t cs = (int64_t)blockIdx.y * nr * nz;  // case of a batch
  extern __shared__ unsigned char smem_raw[];
  taper_dct2_row<T, PER, false>(src + cs, tapered ? tapered + cs : nullptr, dst + cs, rscale, wz,
                                wr, nr, nz, w, fp, tw, e1, e2, blockIdx.x,
                                reinterpret_cast<Cplx<T>*>(smem_raw), nullptr, threadIdx.x,
                                blockDim.x, 0, 0);
}
// One row of the inverse pass (src, dst at the case); buf, tl, tid, nth,
// wtid, wnth and STAGE as for taper_dct2_row.
template <typename T, int PER, bool STAGE>
__device__ __forceinline__ void dct3_row(const T* __restrict__ src, T* __restrict__ dst, int nz,
                                         const FftPlanArgs& fp,
                                         const Cplx<double>* __restrict__ tw,
                                         const Cplx<double>* __restrict__ e1,
                                         const Cplx<double>* __restrict__ e2, int i,
                                         Cplx<T>* buf, Cplx<T>* tl, int tid, int nth, int wtid,
                                         int wnth) {
  const int M = fp.M;
  const T* row = src + (int64_t)i * nz;
  constexpr int NQ = 2 * PER;
  if constexpr (STAGE) stage_twiddles(tl, tw, M, wtid, wnth);
  dct3_pre<T, NQ>(buf, M, e1, e2, [&](int k) { return row[k]; }, tid, nth);
  __syncthreads();
  if constexpr (STAGE)
    fft_lds<T, PER>(buf, fp, static_cast<const Cplx<T>*>(tl), +1, tid, nth);
  else
    fft_lds<T, PER>(buf, fp, tw, +1, tid, nth);
  const T invN = T(1) / T(nz);
  T* out = dst + (int64_t)i * nz;
  constexpr int NJ = 4 * PER;
#pragma unroll
  for (int q = 0; q < NJ; ++q) {
    const int j = tid + q * nth;
    if (j < nz) out[j] = packed_component(buf, dct_dst_index(j, nz)) * invN;
  }
}
template <typename T, int PER>
__global__ void __launch_bounds__(256) dct3_rows_kernel(
    const T* __restrict__ src, T* __restrict__ dst, int nz, FftPlanArgs fp,
    const Cplx<double>* __restrict__ tw, const Cplx<double>* __restrict__ e1,
    const Cplx<double>* __restrict__ e2) {
  const int64_t cs = (int64_t)blockIdx.y * gridDim.x * nz;  // case of a batch
  extern __shared__ unsigned char smem_raw[];
  dct3_row<T, PER, false>(src + cs, dst + cs, nz, fp, tw, e1, e2, blockIdx.x,
                          reinterpret_cast<Cplx<T>*>(smem_raw), nullptr, threadIdx.x, blockDim.x,
                          0, 0);
}

// nz = 2M = 1024 and 2048 row passes: one wavefront per row, FFT in
// registers (axs_fft1024.hpp).  Same inputs, outputs and taper as the
// kernels above.
// Row i of one case (src, tapered, dst at the case), one wavefront, buf of
// kWaveLds<M> entries.
template <typename T, bool TAPER, int M>
__device__ __forceinline__ void dct2_row_wave(
    const T* __restrict__ src, T* __restrict__ tapered, T* __restrict__ dst,
    const double* __restrict__ rscale, const double* __restrict__ wz,
    const double* __restrict__ wr, int nr, int w, const Cplx<double>* __restrict__ tw,
    const Cplx<double>* __restrict__ e1, const Cplx<double>* __restrict__ e2, int i,
    Cplx<T>* buf) {
  constexpr int N = 2 * M;
  const int l = threadIdx.x;
  const int srow = (TAPER && i >= nr - w) ? nr - w : i;
  const T* row = src + (int64_t)srow * N;
  T v[N / 64];
#pragma unroll
  for (int q = 0; q < N / 64; ++q) v[q] = row[l + 64 * q];
  T left = T(0), right = T(0), rowf = T(1);
  if (TAPER) {
    left = row[w - 1];
    right = row[N - w];
    if (i >= nr - w) rowf = T(wr[i - (nr - w)]);
  }
  const T rs = rscale ? T(rscale[i]) : T(1);
  T* out_t = (TAPER && tapered) ? tapered + (int64_t)i * N : nullptr;
  // Makhoul's reorder: element j goes to real slot j / 2 (j even) or
  // N - 1 - (j - 1) / 2 (j odd) of the packed sequence; j = l + 64 q has the
  // parity of l, so the slot is a per-lane base plus or minus 32 q.
  T* zr = reinterpret_cast<T*>(buf);
  const int d0 = (l & 1) ? N - 1 - (l >> 1) : (l >> 1);
  const int dq = (l & 1) ? -32 : 32;
  const bool wide = w > 64;  // the z taper reaches past the first/last 64 columns
#pragma unroll
  for (int q = 0; q < N / 64; ++q) {
    const int j = l + 64 * q;
    T t = v[q];
    if (TAPER) {
      if (q == 0 || q == N / 64 - 1 || wide) {
        if (j < w)
          t = T(wz[j]) * left;
        else if (j >= N - w)
          t = T(wz[w + (j - (N - w))]) * right;
      }
      if (i >= nr - w) t = rowf * t;
      if (i == 0) t = T(0);
      if (out_t) out_t[j] = t;
    }
    zr[d0 + dq * q] = rs * t;
  }
  __syncthreads();
  fft_wave<T, -1, M>(buf, tw);
  T* out = dst + (int64_t)i * N;
  dct2_post_pairs<T, M>(buf, e1, e2, out);
}

template <typename T, bool TAPER, int M>
__global__ void __launch_bounds__(64) dct2_rows_wave_kernel(
    const T* __restrict__ src, T* __restrict__ tapered, T* __restrict__ dst,
    const double* __restrict__ rscale, const double* __restrict__ wz,
    const double* __restrict__ wr, int nr, int w, const Cplx<double>* __restrict__ tw,
    const Cplx<double>* __restrict__ e1, const Cplx<double>* __restrict__ e2, StepFinalize fin) {
  if (fin.partials && blockIdx.x == nr) {  // completes the previous step
    __shared__ double fin_red[256];
    const int cb = blockIdx.y;
    const RigidConfig& cfg = fin.cfgs[cb];
    finalize_pending(fin.partials + (int64_t)cb * fin.count, fin.count,
                     fin.slots + (int64_t)cb * kSlotsPerCase, fin.sc + (int64_t)cb * AXS_RIGID_SCALARS,
                     fin.drag ? fin.drag + (int64_t)cb * cfg.history : nullptr, cfg, fin_red);
    return;
  }
  __shared__ Cplx<T> buf[kWaveLds<M>];
  const int64_t cs = (int64_t)blockIdx.y * nr * (2 * M);  // case of a batch
  dct2_row_wave<T, TAPER, M>(src + cs, tapered ? tapered + cs : nullptr, dst + cs, rscale, wz, wr,
                             nr, w, tw, e1, e2, blockIdx.x, buf);
}

// Row i of one case (src, dst at the case), one wavefront, buf of
// kWaveLds<M> entries.
template <typename T, int M>
__device__ __forceinline__ void dct3_row_wave(const T* __restrict__ src, T* __restrict__ dst,
                                              const Cplx<double>* __restrict__ tw,
                                              const Cplx<double>* __restrict__ e1,
                                              const Cplx<double>* __restrict__ e2, int i,
                                              Cplx<T>* buf) {
  constexpr int N = 2 * M;
  // index pairs (k, M - k), k < M / 2, per lane (the last round partial when
  // 64 does not divide M / 2)
  constexpr int KP = (M / 2 + 63) / 64;
  constexpr bool kFull = (M / 2) % 64 == 0;
  typedef T T2 __attribute__((ext_vector_type(2)));
  static_assert(2 * kWaveLds<M> >= N, "the staged row fits the buffer");
  T* raw = reinterpret_cast<T*>(buf);  // the input row, staged before the pre step
  const int l = threadIdx.x;
  const T2* row = reinterpret_cast<const T2*>(src + (int64_t)i * N);
  // Every global load of the pre step is issued before the first wait: the
  // row once (two values per lane and load) and the twiddles of the pairs.
  T2 v[N / 128];
#pragma unroll
  for (int q = 0; q < N / 128; ++q) v[q] = row[l + 64 * q];
  Cplx<double> q2k[KP], q2m[KP], q1k[KP], q1m[KP];
#pragma unroll
  for (int q = 0; q < KP; ++q) {
    const int k = l + 64 * q;
    if (kFull || k < M / 2) {
      q2k[q] = e2[k];
      q2m[q] = e2[M - k];
      q1k[q] = e1[k];
      q1m[q] = e1[M - k];
    }
  }
  const Cplx<double> q2h = e2[M / 2], q1h = e1[M / 2];
#pragma unroll
  for (int q = 0; q < N / 128; ++q) reinterpret_cast<T2*>(raw)[l + 64 * q] = v[q];
  __syncthreads();
  T xk[KP], xnk[KP], xm[KP], xnm[KP];
#pragma unroll
  for (int q = 0; q < KP; ++q) {
    const int k = l + 64 * q;
    if (kFull || k < M / 2) {
      xk[q] = raw[k];
      xnk[q] = (k == 0) ? T(0) : raw[N - k];
      xm[q] = raw[M - k];
      xnm[q] = raw[M + k];
    }
  }
  const T xh = raw[M / 2], xnh = raw[N - M / 2];
  __syncthreads();  // raw is read above and overwritten by Z below
  // DCT-III pre step (as dct3_pre): Z[k] and Z[M - k] share V_k and V_{M-k}
  // with their roles swapped, so each pair is formed from one set of loads
  // with the arithmetic the single-index form does for each of the two.
#pragma unroll
  for (int q = 0; q < KP; ++q) {
    const int k = l + 64 * q;
    if (!kFull && k >= M / 2) break;
    const Cplx<T> Vk = cmul(Cplx<T>{T(q2k[q].x), T(-q2k[q].y)}, Cplx<T>{xk[q], -xnk[q]});
    const Cplx<T> Vm = cmul(Cplx<T>{T(q2m[q].x), T(-q2m[q].y)}, Cplx<T>{xm[q], -xnm[q]});
    {
      const Cplx<T> A = cadd(Vk, cconj(Vm)), D = csub(Vk, cconj(Vm));
      const Cplx<T> B = cmul(D, Cplx<T>{T(q1k[q].x), T(-q1k[q].y)});
      buf[k] = Cplx<T>{A.x - B.y, A.y + B.x};
    }
    if (k != 0) {
      const Cplx<T> A = cadd(Vm, cconj(Vk)), D = csub(Vm, cconj(Vk));
      const Cplx<T> B = cmul(D, Cplx<T>{T(q1m[q].x), T(-q1m[q].y)});
      buf[M - k] = Cplx<T>{A.x - B.y, A.y + B.x};
    }
  }
  if (l == 0) {  // k = M / 2 is its own partner
    const Cplx<T> Vk = cmul(Cplx<T>{T(q2h.x), T(-q2h.y)}, Cplx<T>{xh, -xnh});
    const Cplx<T> A = cadd(Vk, cconj(Vk)), D = csub(Vk, cconj(Vk));
    const Cplx<T> B = cmul(D, Cplx<T>{T(q1h.x), T(-q1h.y)});
    buf[M / 2] = Cplx<T>{A.x - B.y, A.y + B.x};
  }
  __syncthreads();
  fft_wave<T, +1, M>(buf, tw);
  const T invN = T(1) / T(N);
  T* out = dst + (int64_t)i * N;
  // output j = l + 64 q reads packed slot j / 2 or N - 1 - (j - 1) / 2 (the
  // parity of j is that of l): a per-lane base plus or minus 32 q
  const T* zr = reinterpret_cast<const T*>(buf);
  const int d0 = (l & 1) ? N - 1 - (l >> 1) : (l >> 1);
  const int dq = (l & 1) ? -32 : 32;
#pragma unroll
  for (int q = 0; q < N / 64; ++q) out[l + 64 * q] = zr[d0 + dq * q] * invN;
}

template <typename T, int M>
__global__ void __launch_bounds__(64) dct3_rows_wave_kernel(
    const T* __restrict__ src, T* __restrict__ dst, const Cplx<double>* __restrict__ tw,
    const Cplx<double>* __restrict__ e1, const Cplx<double>* __restrict__ e2) {
  __shared__ Cplx<T> buf[kWaveLds<M>];
  const int64_t cs = (int64_t)blockIdx.y * gridDim.x * (2 * M);  // case of a batch
  dct3_row_wave<T, M>(src + cs, dst + cs, tw, e1, e2, blockIdx.x, buf);
}

// Uneven batches: the rows of every group with a one-wave register transform
// (G.wave: nz = 256, 384, 512, 768, 1024) in one launch, one wavefront per row; block
// b is row b - G.wr0 of its group's cases.
template <typename T>
__global__ void __launch_bounds__(64) taper_dct2_wave_ragged_kernel(
    const T* __restrict__ src, T* __restrict__ dst, const RaggedGroup* __restrict__ grp, int ngrp,
    int w) {
  __shared__ Cplx<T> buf[kF512Lds];
  const int b = blockIdx.x;
  const RaggedGroup& G = grp[ragged_group<&RaggedGroup::wr0>(grp, ngrp, b)];
  const int row = b - G.wr0, cl = row / G.nr, i = row % G.nr;
  const int64_t cs = G.cell0 + (int64_t)cl * G.nr * G.nz;
#define AXS_WAVE_ROW(MM)                                                                     \
  dct2_row_wave<T, true, MM>(src + cs, nullptr, dst + cs, G.rvec, G.kill_wz, G.kill_wr, G.nr, \
                             w, G.tw, G.e1, G.e2, i, buf)
  if (G.nz == 1024)
    AXS_WAVE_ROW(512);
  else if (G.nz == 768)
    AXS_WAVE_ROW(384);
  else if (G.nz == 512)
    AXS_WAVE_ROW(256);
  else if (G.nz == 384)
    AXS_WAVE_ROW(192);
  else
    AXS_WAVE_ROW(128);
#undef AXS_WAVE_ROW
}

template <typename T>
__global__ void __launch_bounds__(64) dct3_wave_ragged_kernel(
    const T* __restrict__ src, T* __restrict__ dst, const RaggedGroup* __restrict__ grp, int ngrp) {
  __shared__ Cplx<T> buf[kF512Lds];
  const int b = blockIdx.x;
  const RaggedGroup& G = grp[ragged_group<&RaggedGroup::wr0>(grp, ngrp, b)];
  const int row = b - G.wr0, cl = row / G.nr, i = row % G.nr;
  const int64_t cs = G.cell0 + (int64_t)cl * G.nr * G.nz;
  if (G.nz == 1024)
    dct3_row_wave<T, 512>(src + cs, dst + cs, G.tw, G.e1, G.e2, i, buf);
  else if (G.nz == 768)
    dct3_row_wave<T, 384>(src + cs, dst + cs, G.tw, G.e1, G.e2, i, buf);
  else if (G.nz == 512)
    dct3_row_wave<T, 256>(src + cs, dst + cs, G.tw, G.e1, G.e2, i, buf);
  else if (G.nz == 384)
    dct3_row_wave<T, 192>(src + cs, dst + cs, G.tw, G.e1, G.e2, i, buf);
  else
    dct3_row_wave<T, 128>(src + cs, dst + cs, G.tw, G.e1, G.e2, i, buf);
}

// Two wavefronts per row (axs_fft1024.hpp, M = 2 HM): nz = 4096 (HM = 1024).
// Same inputs, outputs and taper as the kernels above.  (HM = 512 would
// serve nz = 2048 with two waves per row; it measured slower there than the
// one-wave kernels: 12.3 vs 11.3 us forward, 10.5 vs 9.9 us inverse.)
template <typename T, bool TAPER, int HM>
__global__ void __launch_bounds__(128) dct2_rows_split_kernel(
    const T* __restrict__ src, T* __restrict__ tapered, T* __restrict__ dst,
    const double* __restrict__ rscale, const double* __restrict__ wz,
    const double* __restrict__ wr, int nr, int w, const Cplx<double>* __restrict__ tw,
    const Cplx<double>* __restrict__ e1, const Cplx<double>* __restrict__ e2, StepFinalize fin) {
  constexpr int N = 4 * HM;
  if (fin.partials && blockIdx.x == nr) {  // completes the previous step
    __shared__ double fin_red[256];
    const int cb = blockIdx.y;
    const RigidConfig& cfg = fin.cfgs[cb];
    finalize_pending(fin.partials + (int64_t)cb * fin.count, fin.count,
                     fin.slots + (int64_t)cb * kSlotsPerCase, fin.sc + (int64_t)cb * AXS_RIGID_SCALARS,
                     fin.drag ? fin.drag + (int64_t)cb * cfg.history : nullptr, cfg, fin_red);
    return;
  }
  __shared__ Cplx<T> bufs[2 * kWaveLds<HM>];
  const int t = threadIdx.x;
  const int i = blockIdx.x;
  const int64_t cs = (int64_t)blockIdx.y * nr * N;  // case of a batch
  src += cs;
  dst += cs;
  const int srow = (TAPER && i >= nr - w) ? nr - w : i;
  const T* row = src + (int64_t)srow * N;
  T v[N / 128];
#pragma unroll
  for (int q = 0; q < N / 128; ++q) v[q] = row[t + 128 * q];
  T left = T(0), right = T(0), rowf = T(1);
  if (TAPER) {
    left = row[w - 1];
    right = row[N - w];
    if (i >= nr - w) rowf = T(wr[i - (nr - w)]);
  }
  const T rs = rscale ? T(rscale[i]) : T(1);
  T* out_t = (TAPER && tapered) ? tapered + cs + (int64_t)i * N : nullptr;
  // Makhoul's reorder into the packed sequence, its even and odd complex
  // samples into the two buffers
  T* zr = reinterpret_cast<T*>(bufs);
#pragma unroll
  for (int q = 0; q < N / 128; ++q) {
    const int j = t + 128 * q;
    T x = v[q];
    if (TAPER) {
      if (j < w)
        x = T(wz[j]) * left;
      else if (j >= N - w)
        x = T(wz[w + (j - (N - w))]) * right;
      if (i >= nr - w) x = rowf * x;
      if (i == 0) x = T(0);
      if (out_t) out_t[j] = x;
    }
    const int d = dct_dst_index(j, N);  // real slot of the packed sequence
    zr[2 * split_in_slot<HM>(d >> 1) + (d & 1)] = rs * x;
  }
  __syncthreads();
  fft_wave<T, -1, HM, 2>(bufs + (t >> 6) * kWaveLds<HM>, tw);
  split_join<T, -1, HM>(bufs, tw);
  dct2_post_pairs_split<T, HM>(bufs, e1, e2, dst + (int64_t)i * N);
}

template <typename T, int HM>
__global__ void __launch_bounds__(128) dct3_rows_split_kernel(
    const T* __restrict__ src, T* __restrict__ dst, const Cplx<double>* __restrict__ tw,
    const Cplx<double>* __restrict__ e1, const Cplx<double>* __restrict__ e2) {
  constexpr int M = 2 * HM, N = 4 * HM;
  constexpr int KP = M / 2 / 128;  // index pairs (k, M - k) per thread
  typedef T T2 __attribute__((ext_vector_type(2)));
  __shared__ Cplx<T> bufs[2 * kWaveLds<HM>];
  static_assert(2 * 2 * kWaveLds<HM> >= N, "the staged row fits the two buffers");
  T* raw = reinterpret_cast<T*>(bufs);  // the input row, staged before the pre step
  const int t = threadIdx.x;
  const int i = blockIdx.x;
  const int64_t cs = (int64_t)blockIdx.y * gridDim.x * N;  // case of a batch
  const T2* row = reinterpret_cast<const T2*>(src + cs + (int64_t)i * N);
  T2 v[N / 256];
#pragma unroll
  for (int q = 0; q < N / 256; ++q) v[q] = row[t + 128 * q];
  Cplx<double> q2k[KP], q2m[KP], q1k[KP], q1m[KP];
#pragma unroll
  for (int q = 0; q < KP; ++q) {
    const int k = t + 128 * q;
    q2k[q] = e2[k];
    q2m[q] = e2[M - k];
    q1k[q] = e1[k];
    q1m[q] = e1[M - k];
  }
  const Cplx<double> q2h = e2[M / 2], q1h = e1[M / 2];
#pragma unroll
  for (int q = 0; q < N / 256; ++q) reinterpret_cast<T2*>(raw)[t + 128 * q] = v[q];
  __syncthreads();
  T xk[KP], xnk[KP], xm[KP], xnm[KP];
#pragma unroll
  for (int q = 0; q < KP; ++q) {
    const int k = t + 128 * q;
    xk[q] = raw[k];
    xnk[q] = (k == 0) ? T(0) : raw[N - k];
    xm[q] = raw[M - k];
    xnm[q] = raw[M + k];
  }
  const T xh = raw[M / 2], xnh = raw[N - M / 2];
  __syncthreads();  // raw is read above and overwritten by Z below
  // DCT-III pre step (as dct3_rows_wave_kernel), Z_k to its split slot
#pragma unroll
  for (int q = 0; q < KP; ++q) {
    const int k = t + 128 * q;
    const Cplx<T> Vk = cmul(Cplx<T>{T(q2k[q].x), T(-q2k[q].y)}, Cplx<T>{xk[q], -xnk[q]});
    const Cplx<T> Vm = cmul(Cplx<T>{T(q2m[q].x), T(-q2m[q].y)}, Cplx<T>{xm[q], -xnm[q]});
    {
      const Cplx<T> A = cadd(Vk, cconj(Vm)), D = csub(Vk, cconj(Vm));
      const Cplx<T> B = cmul(D, Cplx<T>{T(q1k[q].x), T(-q1k[q].y)});
      bufs[split_in_slot<HM>(k)] = Cplx<T>{A.x - B.y, A.y + B.x};
    }
    if (k != 0) {
      const Cplx<T> A = cadd(Vm, cconj(Vk)), D = csub(Vm, cconj(Vk));
      const Cplx<T> B = cmul(D, Cplx<T>{T(q1m[q].x), T(-q1m[q].y)});
      bufs[split_in_slot<HM>(M - k)] = Cplx<T>{A.x - B.y, A.y + B.x};
    }
  }
  if (t == 0) {  // k = M / 2 is its own partner
    const Cplx<T> Vk = cmul(Cplx<T>{T(q2h.x), T(-q2h.y)}, Cplx<T>{xh, -xnh});
    const Cplx<T> A = cadd(Vk, cconj(Vk)), D = csub(Vk, cconj(Vk));
    const Cplx<T> B = cmul(D, Cplx<T>{T(q1h.x), T(-q1h.y)});
    bufs[split_in_slot<HM>(M / 2)] = Cplx<T>{A.x - B.y, A.y + B.x};
  }
  __syncthreads();
  fft_wave<T, +1, HM, 2>(bufs + (t >> 6) * kWaveLds<HM>, tw);
  split_join<T, +1, HM>(bufs, tw);
  const T invN = T(1) / T(N);
  T* out = dst + cs + (int64_t)i * N;
  const T* zr = reinterpret_cast<const T*>(bufs);
#pragma unroll
  for (int q = 0; q < N / 128; ++q) {
    const int j = t + 128 * q;
    const int d = dct_dst_index(j, N);
    out[j] = zr[2 * split_out_slot<HM>(d >> 1) + (d & 1)] * invN;
  }
}

// Column pass: for each z-mode k solve the tridiagonal system
//   a_i x_{i-1} + (b_i + lam_k) x_i + c_i x_{i+1} = d_i,   i in [0, nr)
// in place on d (row-major (nr, nz), k contiguous).  Workgroup = K modes x
// P radial segments of L rows; thread (kk, p) owns rows [pL, pL+L) of mode
// k in registers.  Thomas' elimination is split into segment-local
// recurrences joined by prefix scans of affine maps across segments; the
// data-independent pivot recurrence starts each segment from c'_{pL-1},
// tabulated once per plan.  All of a thread's column values are loaded
// before the recurrence starts and the coefficients sit in LDS, so the
// serial chains never wait on global memory.  Workgroups are mapped so that
// neighbouring mode chunks run on the same XCD and share its L2 lines.
// 1 / m to within an ulp: the hardware estimate and two Newton steps (the
// solve is not a correctly rounded quotient anyway; pivots are O(1/dx^2),
// far from the range the IEEE sequence's scaling guards).
__device__ __forceinline__ double recip_nr(double m) {
  double y = __builtin_amdgcn_rcp(m);
  double e = fma(-m, y, 1.0);
  y = fma(y, e, y);
  e = fma(-m, y, 1.0);
  return fma(y, e, y);
}

// FULL: nr = P L and every chunk holds K modes, so no thread or row is idle
// and no bounds test is needed.  CG: the coefficients are read from global
// memory (every mode reads the same rows, so they stay in L1/L2) instead of
// being staged in LDS, which leaves the LDS to the scans and allows more
// modes per workgroup on tall grids.
template <typename T, int L, bool FULL, bool CG>
__device__ __forceinline__ void tridiag_columns_body(
    T* __restrict__ d, const double* __restrict__ a, const double* __restrict__ b,
    const double* __restrict__ c, const double* __restrict__ lam,
    const double* __restrict__ cseg, int nr, int nz, int K, int P, int chunk) {
  extern __shared__ double sm[];
  // coefficient rows are padded by one double per L rows: the threads of a
  // wave read rows p L + l of different segments p in the same instruction,
  // and a stride of L doubles would put them all in one LDS bank
  const int nrp = CG ? 0 : tri_coef_stride(nr, L);
  double* sA = sm;
  double* sB = sm + nrp;
  double* sC = sm + 2 * nrp;
  double* sG0 = sm + 3 * nrp;
  double* sY0 = sG0 + K * P;
  double* sG1 = sY0 + K * P;
  double* sY1 = sG1 + K * P;
  if constexpr (!CG) {
    for (int i = threadIdx.x; i < nr; i += blockDim.x) {
      sA[i + i / L] = a[i];
      sB[i + i / L] = b[i];
      sC[i + i / L] = c[i];
    }
  }
  const int kk = threadIdx.x % K;
  const int p = threadIdx.x / K;
  const int k = chunk * K + kk;
  const bool active = FULL || ((k < nz) && (p < P));
  const int s = p * L;
  const int kc = active ? k : 0;
  T* col = d + ((int64_t)s * nz + kc);  // row s of this thread's mode
  // coefficients of row i = s + l (padded LDS index i + p)
  auto ca = [&](int i) -> double {
    if constexpr (CG) return a[i]; else return sA[i + p];
  };
  auto cb = [&](int i) -> double {
    if constexpr (CG) return b[i]; else return sB[i + p];
  };
  auto cc = [&](int i) -> double {
    if constexpr (CG) return c[i]; else return sC[i + p];
  };

  double x[L];
  double im[L];
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const int i = s + l;
    x[l] = (FULL || (active && i < nr)) ? static_cast<double>(col[(int64_t)l * nz]) : 0.0;
  }
  const double lk = lam[kc];
  double cp = (p == 0 || !active) ? 0.0 : cseg[(size_t)p * nz + kc];
  __syncthreads();
  // -------- forward: local elimination with zero carry-in
  double G = 1.0, y = 0.0;
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const int i = s + l;
    if (FULL || (active && i < nr)) {
      const double ai = ca(i);
      const double m = (cb(i) + lk) - ai * cp;
      const double inv = recip_nr(m);
      im[l] = inv;
      cp = cc(i) * inv;
      y = (x[l] - ai * y) * inv;
      G = G * (-ai * inv);
      x[l] = y;
    } else {
      im[l] = 0.0;
    }
  }
  // inclusive scan over segments of the affine maps D -> G*D + y
  const int me = p * K + kk;
  // threads past K P (a wider ragged launch) only join the barriers; with
  // FULL the workgroup is exactly K P threads
  const bool seg = FULL || p < P;
  if (seg) {
    sG0[me] = G;
    sY0[me] = y;
  }
  __syncthreads();
  double* inG = sG0; double* inY = sY0; double* outG = sG1; double* outY = sY1;
  for (int off = 1; off < P; off <<= 1) {
    if (seg) {
      double g = inG[me], yy = inY[me];
      if (p >= off) {
        const int o = (p - off) * K + kk;
        // compose: this o (earlier) then me: D -> g*(G_o D + y_o) + yy
        yy = g * inY[o] + yy;
        g = g * inG[o];
      }
      outG[me] = g;
      outY[me] = yy;
    }
    __syncthreads();
    double* t;
    t = inG; inG = outG; outG = t;
    t = inY; inY = outY; outY = t;
  }
  const double Din = (p == 0 || !seg) ? 0.0 : inY[(p - 1) * K + kk];
  __syncthreads();
  // fix-up: d'_i = y_i + G_i * Din
  {
    double g = 1.0;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const int i = s + l;
      if (FULL || (active && i < nr)) {
        g = g * (-ca(i) * im[l]);
        x[l] = x[l] + g * Din;
      }
    }
  }
  // -------- backward: local substitution with zero carry-in from above
  double H = 1.0, z = 0.0;
#pragma unroll
  for (int l = L - 1; l >= 0; --l) {
    const int i = s + l;
    if (FULL || (active && i < nr)) {
      const double cpi = cc(i) * im[l];
      z = x[l] - cpi * z;
      H = H * (-cpi);
      x[l] = z;
    }
  }
  // suffix scan over segments (later segment feeds the earlier one)
  if (seg) {
    sG0[me] = H;
    sY0[me] = z;
  }
  __syncthreads();
  inG = sG0; inY = sY0; outG = sG1; outY = sY1;
  for (int off = 1; off < P; off <<= 1) {
    if (seg) {
      double h = inG[me], zz = inY[me];
      if (p + off < P) {
        const int o = (p + off) * K + kk;
        zz = h * inY[o] + zz;
        h = h * inG[o];
      }
      outG[me] = h;
      outY[me] = zz;
    }
    __syncthreads();
    double* t;
    t = inG; inG = outG; outG = t;
    t = inY; inY = outY; outY = t;
  }
  const double Xin = (p >= P - 1) ? 0.0 : inY[(p + 1) * K + kk];
  // fix-up and store: x_i = z_i + H_i * Xin, H_i = prod_{j=i}^{e-1} (-c'_j)
  {
    double h = 1.0;
#pragma unroll
    for (int l = L - 1; l >= 0; --l) {
      const int i = s + l;
      if (FULL || (active && i < nr)) {
        h = h * (-(cc(i) * im[l]));
        col[(int64_t)l * nz] = static_cast<T>(x[l] + h * Xin);
      }
    }
  }
}
template <typename T, int L, bool FULL, bool CG>
__global__ void __launch_bounds__(1024) tridiag_columns_kernel(
    T* __restrict__ d, const double* __restrict__ a, const double* __restrict__ b,
    const double* __restrict__ c, const double* __restrict__ lam,
    const double* __restrict__ cseg, int nr, int nz, int K, int P) {
  // XCD-aware chunk order: block b runs on XCD b % 8; give each XCD a
  // contiguous range of mode chunks.
  const int nb = gridDim.x;
  int chunk = blockIdx.x;
  if (nb % 8 == 0) chunk = (blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8;
  tridiag_columns_body<T, L, FULL, CG>(d + (int64_t)blockIdx.y * nr * nz, a, b, c, lam, cseg, nr,
                                       nz, K, P, chunk);
}

// ------------------------------------------------------------ uneven batches

template <typename T>
__global__ void __launch_bounds__(256) taper_dct2_ragged_kernel(
    const T* __restrict__ src, T* __restrict__ dst, const RaggedGroup* __restrict__ grp, int ngrp,
    int ncases, int w, StepFinalize fin) {
  int b = blockIdx.x;
  if (fin.partials) {
    if (b < ncases) {  // completes case b's previous step
      __shared__ double fin_red[256];
      const RaggedGroup& G = grp[ragged_group<&RaggedGroup::case0>(grp, ngrp, b)];
      const RigidConfig& cfg = fin.cfgs[b];
      finalize_pending(fin.partials + G.part0 + (int64_t)(b - G.case0) * G.ftiles, G.ftiles,
                       fin.slots + (int64_t)b * kSlotsPerCase,
                       fin.sc + (int64_t)b * AXS_RIGID_SCALARS,
                       fin.drag ? fin.drag + (int64_t)b * cfg.history : nullptr, cfg, fin_red);
      return;
    }
    b -= ncases;
  }
  const RaggedGroup& G = grp[ragged_group<&RaggedGroup::rb0>(grp, ngrp, b)];
  const int sub = threadIdx.x / G.bs;  // row of this workgroup
  const int row = (b - G.rb0) * G.rpb + sub, cl = row / G.nr, i = row % G.nr;
  const int64_t cs = G.cell0 + (int64_t)cl * G.nr * G.nz;
  extern __shared__ unsigned char smem_raw[];
  Cplx<T>* lds = reinterpret_cast<Cplx<T>*>(smem_raw);
  taper_dct2_row<T, 1, true>(src + cs, nullptr, dst + cs, G.rvec, G.kill_wz, G.kill_wr, G.nr, G.nz, w,
                       G.fp, G.tw, G.e1, G.e2, i, lds + sub * G.fp.M, lds + G.rpb * G.fp.M,
                       threadIdx.x % G.bs, G.bs, threadIdx.x, blockDim.x);
}

template <typename T>
__global__ void __launch_bounds__(256) dct3_ragged_kernel(const T* __restrict__ src,
                                                          T* __restrict__ dst,
                                                          const RaggedGroup* __restrict__ grp,
                                                          int ngrp) {
  const int b = blockIdx.x;
  const RaggedGroup& G = grp[ragged_group<&RaggedGroup::rb0>(grp, ngrp, b)];
  const int sub = threadIdx.x / G.bs;
  const int row = (b - G.rb0) * G.rpb + sub, cl = row / G.nr, i = row % G.nr;
  const int64_t cs = G.cell0 + (int64_t)cl * G.nr * G.nz;
  extern __shared__ unsigned char smem_raw[];
  Cplx<T>* lds = reinterpret_cast<Cplx<T>*>(smem_raw);
  dct3_row<T, 1, true>(src + cs, dst + cs, G.nz, G.fp, G.tw, G.e1, G.e2, i, lds + sub * G.fp.M,
                 lds + G.rpb * G.fp.M, threadIdx.x % G.bs, G.bs, threadIdx.x, blockDim.x);
}

template <typename T, int L>
__global__ void __launch_bounds__(1024) tridiag_ragged_kernel(T* __restrict__ d,
                                                             const RaggedGroup* __restrict__ grp,
                                                             int ngrp) {
  const int b = blockIdx.x;
  const RaggedGroup& G = grp[ragged_group<&RaggedGroup::tri0>(grp, ngrp, b)];
  const int q = b - G.tri0, cl = q / G.tchunks, chunk = q % G.tchunks;
  tridiag_columns_body<T, L, false, false>(d + G.cell0 + (int64_t)cl * G.nr * G.nz, G.a, G.b,
                                           G.c, G.lam, G.cseg, G.nr, G.nz, G.K, G.P, chunk);
}

// Soft batches: mode chunk blockIdx.x of case blockIdx.y, segments of
// kSoftBatchTriL rows (every case's plan is made for that L); the spectrum
// in the batch's storage type, the elimination in FP64.
template <typename B>
__global__ void __launch_bounds__(1024) soft_batch_tridiag_kernel(
    const SoftCaseT<B>* __restrict__ cs) {
  const SoftCaseT<B>& C = cs[blockIdx.y];
  if ((int)blockIdx.x * C.tK >= C.nz) return;
  tridiag_columns_body<B, kSoftBatchTriL, false, false>(C.spec, C.ta, C.tb, C.tc, C.tlam, C.tcseg,
                                                        C.nr, C.nz, C.tK, C.tP, blockIdx.x);
}

template <typename B>
hipError_t launch_soft_batch_tridiag(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                     hipStream_t s) {
  soft_batch_tridiag_kernel<B><<<dim3(d.max_chunks, d.ncase), 1024, d.tri_lds, s>>>(cases);
  return hipGetLastError();
}
template hipError_t launch_soft_batch_tridiag<double>(const SoftCaseT<double>*,
                                                      const SoftBatchDims&, hipStream_t);
template hipError_t launch_soft_batch_tridiag<float>(const SoftCaseT<float>*,
                                                     const SoftBatchDims&, hipStream_t);

// ------------------------------------------------------------------ launchers

// Half-lengths with a one-wave register transform (axs_fft1024.hpp).
bool wave_rows(int M) {
  return M == 128 || M == 192 || M == 256 || M == 384 || M == 512 || M == 1024;
}

static int fft_block(int M) {
  // threads so that every stage needs at most PER butterflies per thread
  return std::min(256, std::max(64, M / 2));
}

// A plan whose row tables were never built (a radial-only plan) refuses
// the row passes instead of handing their kernels null tables.
static bool has_row_tables(const StokesPlan& pl) {
  return pl.general ? pl.any != nullptr : (pl.tw && pl.e1 && pl.e2);
}

template <typename T>
hipError_t launch_dct2_rows(const StokesPlan& pl, const T* src, T* dst, const double* rscale,
                            int rows, hipStream_t s) {
  if (!has_row_tables(pl)) return hipErrorInvalidValue;
  if (pl.general)
    return launch_any_dct2_rows<T>(pl, src, nullptr, dst, rscale, nullptr, nullptr, 0, false, 1,
                                   nullptr, s);
  const int M = pl.nz / 2;
  const int bs = fft_block(M);
  // worst stage is radix 2 (M/2 butterflies); 2*per*bs >= M covers the row loads
  const int per = std::max(1, (M / 2 + bs - 1) / bs);
  const size_t shm = sizeof(Cplx<T>) * M;
  if (wave_rows(M)) {
#define AXS_DCT2_WAVE(MM)                                                                     \
  dct2_rows_wave_kernel<T, false, MM><<<rows, 64, 0, s>>>(                                     \
      src, nullptr, dst, rscale, nullptr, nullptr, rows, 0, pl.tw, pl.e1, pl.e2, StepFinalize{})
    if (M == 1024) AXS_DCT2_WAVE(1024);
    else if (M == 512) AXS_DCT2_WAVE(512);
    else if (M == 384) AXS_DCT2_WAVE(384);
    else if (M == 256) AXS_DCT2_WAVE(256);
    else if (M == 192) AXS_DCT2_WAVE(192);
    else AXS_DCT2_WAVE(128);
#undef AXS_DCT2_WAVE
    return hipGetLastError();
  }
  if (M == kF2048M) {
    dct2_rows_split_kernel<T, false, 1024><<<rows, 128, 0, s>>>(
        src, nullptr, dst, rscale, nullptr, nullptr, rows, 0, pl.tw, pl.e1, pl.e2, StepFinalize{});
    return hipGetLastError();
  }
  if (per <= 1)
    dct2_rows_kernel<T, 1><<<rows, bs, shm, s>>>(src, dst, rscale, pl.nz, pl.fft, pl.tw, pl.e1, pl.e2);
  else if (per <= 2)
    dct2_rows_kernel<T, 2><<<rows, bs, shm, s>>>(src, dst, rscale, pl.nz, pl.fft, pl.tw, pl.e1, pl.e2);
  else if (per <= 4)
    dct2_rows_kernel<T, 4><<<rows, bs, shm, s>>>(src, dst, rscale, pl.nz, pl.fft, pl.tw, pl.e1, pl.e2);
  else
    dct2_rows_kernel<T, 8><<<rows, bs, shm, s>>>(src, dst, rscale, pl.nz, pl.fft, pl.tw, pl.e1, pl.e2);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_taper_dct2_rows(const StokesPlan& pl, const T* src, T* tapered, T* dst,
                                  const double* wz, const double* wr, int w, hipStream_t s,
                                  int batch, const StepFinalize* fin) {
  if (!has_row_tables(pl)) return hipErrorInvalidValue;
  if (pl.general)
    return launch_any_dct2_rows<T>(pl, src, tapered, dst, pl.rvec, wz, wr, w, true, batch, fin,
                                   s);
  const StepFinalize f = fin ? *fin : StepFinalize{};
  const int gx = pl.nr + (f.partials ? 1 : 0);
  const int M = pl.nz / 2;
  const int bs = fft_block(M);
  const int per = std::max(1, (M / 2 + bs - 1) / bs);
  const size_t shm = sizeof(Cplx<T>) * M;
  if (wave_rows(M)) {
#define AXS_TAPER_WAVE(MM)                                                                    \
  dct2_rows_wave_kernel<T, true, MM><<<dim3(gx, batch), 64, 0, s>>>(                           \
      src, tapered, dst, pl.rvec, wz, wr, pl.nr, w, pl.tw, pl.e1, pl.e2, f)
    if (M == 1024) AXS_TAPER_WAVE(1024);
    else if (M == 512) AXS_TAPER_WAVE(512);
    else if (M == 384) AXS_TAPER_WAVE(384);
    else if (M == 256) AXS_TAPER_WAVE(256);
    else if (M == 192) AXS_TAPER_WAVE(192);
    else AXS_TAPER_WAVE(128);
#undef AXS_TAPER_WAVE
    return hipGetLastError();
  }
  if (M == kF2048M) {
    dct2_rows_split_kernel<T, true, 1024><<<dim3(gx, batch), 128, 0, s>>>(
        src, tapered, dst, pl.rvec, wz, wr, pl.nr, w, pl.tw, pl.e1, pl.e2, f);
    return hipGetLastError();
  }
#define AXS_TAPER_LAUNCH(P)                                                                    \
  taper_dct2_rows_kernel<T, P><<<dim3(gx, batch), bs, shm, s>>>(                              \
      src, tapered, dst, pl.rvec, wz, wr, pl.nr, pl.nz, w, pl.fft, pl.tw, pl.e1, pl.e2, f)
  if (per <= 1)
    AXS_TAPER_LAUNCH(1);
  else if (per <= 2)
    AXS_TAPER_LAUNCH(2);
  else if (per <= 4)
    AXS_TAPER_LAUNCH(4);
  else
    AXS_TAPER_LAUNCH(8);
#undef AXS_TAPER_LAUNCH
  return hipGetLastError();
}

template <typename T>
hipError_t launch_dct3_rows(const StokesPlan& pl, const T* src, T* dst, int rows, hipStream_t s,
                            int batch) {
  if (!has_row_tables(pl)) return hipErrorInvalidValue;
  if (pl.general) return launch_any_dct3_rows<T>(pl, src, dst, rows, batch, s);
  const int M = pl.nz / 2;
  const int bs = fft_block(M);
  const int per = std::max(1, (M / 2 + bs - 1) / bs);
  const size_t shm = sizeof(Cplx<T>) * M;
  const dim3 grid(rows, batch);
  if (wave_rows(M)) {
#define AXS_DCT3_WAVE(MM) \
  dct3_rows_wave_kernel<T, MM><<<grid, 64, 0, s>>>(src, dst, pl.tw, pl.e1, pl.e2)
    if (M == 1024) AXS_DCT3_WAVE(1024);
    else if (M == 512) AXS_DCT3_WAVE(512);
    else if (M == 384) AXS_DCT3_WAVE(384);
    else if (M == 256) AXS_DCT3_WAVE(256);
    else if (M == 192) AXS_DCT3_WAVE(192);
    else AXS_DCT3_WAVE(128);
#undef AXS_DCT3_WAVE
    return hipGetLastError();
  }
  if (M == kF2048M) {
    dct3_rows_split_kernel<T, 1024><<<grid, 128, 0, s>>>(src, dst, pl.tw, pl.e1, pl.e2);
    return hipGetLastError();
  }
  if (per <= 1)
    dct3_rows_kernel<T, 1><<<grid, bs, shm, s>>>(src, dst, pl.nz, pl.fft, pl.tw, pl.e1, pl.e2);
  else if (per <= 2)
    dct3_rows_kernel<T, 2><<<grid, bs, shm, s>>>(src, dst, pl.nz, pl.fft, pl.tw, pl.e1, pl.e2);
  else if (per <= 4)
    dct3_rows_kernel<T, 4><<<grid, bs, shm, s>>>(src, dst, pl.nz, pl.fft, pl.tw, pl.e1, pl.e2);
  else
    dct3_rows_kernel<T, 8><<<grid, bs, shm, s>>>(src, dst, pl.nz, pl.fft, pl.tw, pl.e1, pl.e2);
  return hipGetLastError();
}

hipError_t launch_rfft_rows(const StokesPlan& pl, const double* src, double* dst,
                            const double* rscale, int rows, hipStream_t s, bool inverse) {
  if (pl.general) return launch_any_rfft_rows(pl, src, dst, rscale, rows, inverse, s);
  const int M = pl.nz / 2;
  const int bs = fft_block(M);
  const int per = std::max(1, (M / 2 + bs - 1) / bs);
  const size_t shm = sizeof(Cplx<double>) * M;
#define AXS_RFFT(P)                                                                           \
  do {                                                                                        \
    if (inverse)                                                                              \
      irfft_rows_kernel<P><<<rows, bs, shm, s>>>(src, dst, pl.nz, pl.fft, pl.tw, pl.e1);     \
    else                                                                                      \
      rfft_rows_kernel<P><<<rows, bs, shm, s>>>(src, dst, rscale, pl.nz, pl.fft, pl.tw, pl.e1); \
  } while (0)
  if (per <= 1)
    AXS_RFFT(1);
  else if (per <= 2)
    AXS_RFFT(2);
  else if (per <= 4)
    AXS_RFFT(4);
  else
    AXS_RFFT(8);
#undef AXS_RFFT
  return hipGetLastError();
}

// One tridiagonal system of nr rows held contiguously in d (segments of 16
// rows, one thread each, coefficients read from global memory).
__global__ void __launch_bounds__(256) tridiag_one_kernel(
    double* __restrict__ d, const double* __restrict__ a, const double* __restrict__ b,
    const double* __restrict__ c, const double* __restrict__ lam,
    const double* __restrict__ cseg, int nr, int P) {
  tridiag_columns_body<double, 16, false, true>(d, a, b, c, lam, cseg, nr, 1, 1, P, 0);
}

hipError_t launch_tridiag_one(double* d, const double* a, const double* b, const double* c,
                              const double* lam, const double* cseg, int nr, hipStream_t s) {
  const int P = (nr + 15) / 16;
  if (P > 256) return hipErrorInvalidValue;
  tridiag_one_kernel<<<1, P, tri_lds_bytes(nr, 16, 1, P, true), s>>>(d, a, b, c, lam, cseg, nr,
                                                                      P);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_tridiag(const StokesPlan& pl, T* d, hipStream_t s, int batch) {
  const int K = pl.tri_K, P = pl.tri_P;
  const dim3 blocks(ceil_div(pl.nz, K), batch);
  const bool full = (pl.nr % pl.tri_L == 0) && (pl.nz % K == 0);
  const bool cg = pl.tri_G;
  const size_t shm = tri_lds_bytes(pl.nr, pl.tri_L, K, P, cg);
#define AXS_TRI_ONE(LL, F, G)                                                                    \
  tridiag_columns_kernel<T, LL, F, G><<<blocks, K * P, shm, s>>>(d, pl.a, pl.b, pl.c, pl.lam,    \
                                                                 pl.cseg, pl.nr, pl.nz, K, P)
#define AXS_TRI(LL)                                                                              \
  do {                                                                                           \
    if (full && cg)                                                                              \
      AXS_TRI_ONE(LL, true, true);                                                               \
    else if (full)                                                                               \
      AXS_TRI_ONE(LL, true, false);                                                              \
    else if (cg)                                                                                 \
      AXS_TRI_ONE(LL, false, true);                                                              \
    else                                                                                         \
      AXS_TRI_ONE(LL, false, false);                                                             \
  } while (0)
  switch (pl.tri_L) {
    case 4:
      AXS_TRI(4);
      break;
    case 8:
      AXS_TRI(8);
      break;
    case 16:
      AXS_TRI(16);
      break;
    case 32:
      AXS_TRI(32);
      break;
    default:
      return hipErrorInvalidValue;
  }
#undef AXS_TRI_ONE
#undef AXS_TRI
  return hipGetLastError();
}

template <typename T>
hipError_t launch_taper_dct2_ragged(const T* src, T* dst, const RaggedGroup* grp, int ngrp,
                                    int ncases, int row_blocks, int w, int lds_cplx,
                                    const StepFinalize* fin, hipStream_t s) {
  const StepFinalize f = fin ? *fin : StepFinalize{};
  const int blocks = row_blocks + (f.partials ? ncases : 0);
  if (blocks == 0) return hipSuccess;  // every group runs its rows elsewhere
  taper_dct2_ragged_kernel<T><<<blocks, 256, sizeof(Cplx<T>) * lds_cplx, s>>>(src, dst, grp, ngrp,
                                                                              ncases, w, f);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_dct3_ragged(const T* src, T* dst, const RaggedGroup* grp, int ngrp,
                              int row_blocks, int lds_cplx, hipStream_t s) {
  dct3_ragged_kernel<T><<<row_blocks, 256, sizeof(Cplx<T>) * lds_cplx, s>>>(src, dst, grp, ngrp);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_wave_rows_ragged(const T* src, T* dst, const RaggedGroup* grp, int ngrp,
                                   int wave_rows, int w, bool forward, hipStream_t s) {
  if (wave_rows == 0) return hipSuccess;
  if (forward)
    taper_dct2_wave_ragged_kernel<T><<<wave_rows, 64, 0, s>>>(src, dst, grp, ngrp, w);
  else
    dct3_wave_ragged_kernel<T><<<wave_rows, 64, 0, s>>>(src, dst, grp, ngrp);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_tridiag_ragged(T* d, const RaggedGroup* grp, int ngrp, int blocks, size_t shm,
                                 int L, hipStream_t s) {
  if (L == 4)
    tridiag_ragged_kernel<T, 4><<<blocks, 1024, shm, s>>>(d, grp, ngrp);
  else if (L == 8)
    tridiag_ragged_kernel<T, 8><<<blocks, 1024, shm, s>>>(d, grp, ngrp);
  else if (L == 16)
    tridiag_ragged_kernel<T, 16><<<blocks, 1024, shm, s>>>(d, grp, ngrp);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

#define AXS_RAGGED_INST(T)                                                                      \
  template hipError_t launch_taper_dct2_ragged<T>(const T*, T*, const RaggedGroup*, int, int, int, \
                                                  int, int, const StepFinalize*, hipStream_t);    \
  template hipError_t launch_dct3_ragged<T>(const T*, T*, const RaggedGroup*, int, int, int,      \
                                            hipStream_t);                                         \
  template hipError_t launch_wave_rows_ragged<T>(const T*, T*, const RaggedGroup*, int, int, int, \
                                                 bool, hipStream_t);                              \
  template hipError_t launch_tridiag_ragged<T>(T*, const RaggedGroup*, int, int, size_t, int,     \
                                               hipStream_t);
AXS_RAGGED_INST(double)
AXS_RAGGED_INST(float)
#undef AXS_RAGGED_INST
template hipError_t launch_dct2_rows<double>(const StokesPlan&, const double*, double*, const double*, int, hipStream_t);
template hipError_t launch_dct3_rows<double>(const StokesPlan&, const double*, double*, int, hipStream_t, int);
template hipError_t launch_tridiag<double>(const StokesPlan&, double*, hipStream_t, int);
template hipError_t launch_dct2_rows<float>(const StokesPlan&, const float*, float*, const double*, int, hipStream_t);
template hipError_t launch_dct3_rows<float>(const StokesPlan&, const float*, float*, int, hipStream_t, int);
template hipError_t launch_tridiag<float>(const StokesPlan&, float*, hipStream_t, int);
template hipError_t launch_taper_dct2_rows<double>(const StokesPlan&, const double*, double*,
                                                   double*, const double*, const double*, int,
                                                   hipStream_t, int, const StepFinalize*);
template hipError_t launch_taper_dct2_rows<float>(const StokesPlan&, const float*, float*, float*,
                                                  const double*, const double*, int, hipStream_t,
                                                  int, const StepFinalize*);

}  // namespace axs

using namespace axs;

struct axs_stokes_plan {
  StokesPlan impl;
};

extern "C" {

axs_status axs_stokes_plan_create(int nr, int nz, double dx, int bc_type,
                                  axs_stokes_plan** plan_out) {
  AXS_CHECK_ARG(plan_out, "null plan_out");
  *plan_out = nullptr;
  AXS_CHECK_ARG(bc_type == AXS_BC_NEUMANN_Z_AND_R || bc_type == AXS_BC_NEUMANN_R_PERIODIC_Z ||
                    bc_type == AXS_BC_DIRICHLET_R_PERIODIC_Z,
                "unknown bc_type");
  AXS_CHECK_ARG(nr >= 3 && nz >= 4, "need nr >= 3 and nz >= 4");
  AXS_CHECK_ARG(nr <= 4096, "nr > 4096: too tall for the column pass");
  AXS_CHECK_ARG(dx > 0.0, "dx must be positive");
  if (!stokes_nz_supported(nz)) {
    set_error("nz too large for the row transforms (nz <= 8192 with nz/2 2,3,5-smooth, "
              "else nz <= 4096)");
    return AXS_ERR_UNSUPPORTED;
  }
  axs_stokes_plan* p = new (std::nothrow) axs_stokes_plan();
  if (!p) {
    set_error("out of host memory");
    return AXS_ERR_ALLOC;
  }
  hipError_t e = p->impl.init(nr, nz, dx, bc_type);
  if (e != hipSuccess) {
    set_error(std::string("stokes plan: ") + hipGetErrorString(e));
    delete p;
    return AXS_ERR_HIP;
  }
  *plan_out = p;
  return AXS_OK;
}

axs_status axs_stokes_plan_destroy(axs_stokes_plan* plan) {
  delete plan;
  return AXS_OK;
}

axs_status axs_stokes_solve_f64(axs_stokes_plan* plan, double* psi, const double* rhs,
                                void* stream) {
  AXS_CHECK_ARG(plan && psi && rhs, "null pointer");
  StokesPlan& pl = plan->impl;
  hipStream_t s = as_stream(stream);
  if (pl.bc != AXS_BC_NEUMANN_Z_AND_R) {  // periodic z: real DFT rows
    AXS_HIP(launch_rfft_rows(pl, rhs, pl.spec, pl.rvec, pl.nr, s, false));
    AXS_HIP(launch_tridiag<double>(pl, pl.spec, s));
    AXS_HIP(launch_rfft_rows(pl, pl.spec, psi, nullptr, pl.nr, s, true));
    return AXS_OK;
  }
  AXS_HIP(launch_dct2_rows<double>(pl, rhs, pl.spec, pl.rvec, pl.nr, s));
  AXS_HIP(launch_tridiag<double>(pl, pl.spec, s));
  AXS_HIP(launch_dct3_rows<double>(pl, pl.spec, psi, pl.nr, s));
  return AXS_OK;
}

}  // extern "C"
