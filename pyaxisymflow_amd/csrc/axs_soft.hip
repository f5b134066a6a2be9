// Device-resident timestep of the soft-body driver (row A16, soft case):
// examples/SoftSphereInTaylorGreenVortex/soft_sphere_in_taylor_green_vortex.py:101-213.
// An elastic sphere tracked by its reference map (eta1, eta2) and level set
// phi in an axisymmetric vortex; no Brinkman term.  One step, in the
// reference's order:
//   taper_dct2_rows   kill taper (kill_boundary_vorticity_sine.py:4-27) into
//                     work, DCT-II of r * omega
//   tridiag_columns   radial tridiagonal per z-mode
//   dct3_rows         -> psi
//   velocity          u_z, u_r from psi (compute_velocity_from_psi.py:4-17)
//   velocity_max      max(|u_z| + |u_r|) into spread atomicMax slots
//   dt                min(CFL dx / (max + eps), CFL dx / sqrt(G / rho_f),
//                     0.9 dx^2 / 4 / nu) (:139-143), derived scalars
//   refmap_eno3       eta advected (advect_refmap_via_eno3.py:8-54)
//   pin_phi           phi = sphere of eta where phi > -3 dx (:152-156)
//   reinit            narrow-band signed distance, masked cells keep phi
//                     (:159-162; axs_reinit.hip)
//   lsq_mirror        chi = smooth Heaviside(phi) (:165-166); doubled-grid
//                     phi, eta, flags; the sweep counters cleared
//                     (extrapolate_eta_using_least_squares_unb.py:19-24,
//                     extrapolate_using_least_squares.py:33-35)
//   lsq_sweeps        least-squares extrapolation sweeps until a sweep
//                     fits nothing, over a list of the cells to fill made
//                     once per step; each sweep is one launch reading one
//                     buffer set and writing the other, and returns at
//                     once after the loop has ended; a loop longer than
//                     the queued sweeps is finished by one workgroup
//   lsq_back          eta = doubled eta's physical half (:29-30), from
//                     the buffer set holding the result
//   solid_sigma       stresses (solid_sigma.py:4-29), blended by chi as they
//                     are stored (:197-199)
//   div_tau           omega += dt curl(div sigma) (div_tau.py:4-34)
//   advect            ENO3 conservative (advect_vorticity_via_eno3.py:8-44)
//   diffuse           RK2 (diffusion_RK2.py:4-44) -> home buffer
//   finalize          t += dt, it += 1, LS convergence record
// FP64 throughout.  Every scalar stays on the device; the step is captured
// into a hipGraph (fills are kernels, no memset nodes).
#include <algorithm>
#include <new>
#include <type_traits>
#include <vector>

#include "axs_eno3.hpp"
#include "axs_fused.hpp"
#include "axs_lsq.hpp"
#include "axs_reinit.hpp"
#include "axs_softbatch.hpp"
#include "axs_stencil.hpp"
#include "axs_stokes.hpp"

namespace axs {

static_assert(kSsCount <= AXS_RIGID_SCALARS, "soft scalar slots exceed the array");

struct SoftConfig {
  int nr, nz, kill_width, lsq_sweeps, reinit_iterations, lsq_grow;
  double dx, nu, G, rho_f, CFL, eps, r_ball, z_cm, r_cm, moll_zone, extrap_zone, reinit_band;
  // streaming (examples/SoftSphereStreaming/soft_sphere_streaming.py:128-264):
  // the vorticity is advected before the solid stress acts, a rigid core
  // of radius tether_rad at (z_cm + e_amp r_ball sin(omega t), r_cm) is
  // penalised (brink_lam) towards U0 cos(omega t) after it, and dt is
  // clipped to the end of the cycle (freq_limit) and to t_end
  int streaming;
  double brink_lam, U0, omega, e_amp, tether_rad, freq_limit, t_end;
};

// ------------------------------------------------------------------ kernels

__global__ void soft_velocity_max_kernel(const double* __restrict__ uz,
                                         const double* __restrict__ ur, int64_t n,
                                         double* __restrict__ slots) {
  __shared__ double wm[4];
  double m = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    m = max_nan(m, fabs(uz[k]) + fabs(ur[k]));
  for (int off = 32; off > 0; off >>= 1) m = max_nan(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = max_nan(max_nan(wm[0], wm[1]), max_nan(wm[2], wm[3]));
    atomicMax(reinterpret_cast<unsigned long long*>(slots + (blockIdx.x % kMaxSlots) * kSlotStride),
              static_cast<unsigned long long>(__double_as_longlong(m)));
  }
}

// One block of 64: the dt rule, derived scalars, cleared slots.
__device__ __forceinline__ void soft_dt_body(double* __restrict__ sc, double* __restrict__ slots,
                                             double CFL, double dx, double eps, double G,
                                             double rho_f, double nu) {
  double m = __longlong_as_double(static_cast<long long>(
      __hip_atomic_load(reinterpret_cast<unsigned long long*>(slots + threadIdx.x * kSlotStride),
                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
  for (int off = 32; off > 0; off >>= 1) m = max_nan(m, __shfl_xor(m, off));
  slots[threadIdx.x * kSlotStride] = 0.0;
  if (threadIdx.x != 0) return;
  double dt = CFL * dx / (m + eps);
  const double d2 = CFL * dx / sqrt(G / rho_f);
  const double d3 = 0.9 * (dx * dx) / 4.0 / nu;  // 0.9 * dx**2 / 4 / nu
  if (d2 < dt) dt = d2;
  if (d3 < dt) dt = d3;
  sc[kSsMaxU] = m;
  sc[kSsDt] = dt;
  sc[kSsDtByDx] = dt / dx;
  sc[kSsHalfNuDt] = 0.5 * nu * dt;
  sc[kSsNuDt] = nu * dt;
}

__global__ void soft_dt_kernel(double* __restrict__ sc, double* __restrict__ slots,
                               SoftConfig cfg) {
  if (cfg.streaming && threadIdx.x == 0 && sc[kSsFreqTimer] >= cfg.freq_limit)
    sc[kSsFreqTimer] = 0.0;  // :128-161 (the cycle's plot resets the timer)
  soft_dt_body(sc, slots, cfg.CFL, cfg.dx, cfg.eps, cfg.G, cfg.rho_f, cfg.nu);
  if (!cfg.streaming || threadIdx.x != 0) return;
  // :168-171
  double dt = sc[kSsDt];
  if (sc[kSsFreqTimer] + dt > cfg.freq_limit) dt = cfg.freq_limit - sc[kSsFreqTimer];
  if (sc[kSsT] + dt > cfg.t_end) dt = cfg.t_end - sc[kSsT];
  sc[kSsDt] = dt;
  sc[kSsDtByDx] = dt / cfg.dx;
  sc[kSsHalfNuDt] = 0.5 * cfg.nu * dt;
  sc[kSsNuDt] = cfg.nu * dt;
  // :243 and :250-260, at the step's start time
  const double t = sc[kSsT];
  sc[kSsZtether] = cfg.z_cm + cfg.e_amp * cfg.r_ball * sin(cfg.omega * t);
  sc[kSsUwall] = cfg.U0 * cos(cfg.omega * t);
}

// The tethered core (:243-264): its mask smooth_Heaviside(tether_rad -
// |(z, r) - (Z_tether, r_cm)|, moll_zone); u_upen = u, u = Brinkman(u_upen)
// towards (U_wall, 0), du = u - u_upen.
__global__ void soft_tether_kernel(double* __restrict__ uz, double* __restrict__ ur,
                                   double* __restrict__ duz, double* __restrict__ dur,
                                   const double* __restrict__ r, const double* __restrict__ z,
                                   const double* __restrict__ sc, SoftConfig cfg) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= cfg.nz) return;
  const int64_t c = (int64_t)i * cfg.nz + j;
  const double a = z[j] - sc[kSsZtether], b = r[i] - cfg.r_cm;
  const double phi = -sqrt(a * a + b * b) + cfg.tether_rad;
  const double chi = smooth_heaviside(phi, cfg.moll_zone);
  const double ldc = (cfg.brink_lam * sc[kSsDt]) * chi;
  const double den = 1.0 + ldc;
  const double gz = uz[c], gr = ur[c];
  const double pz = (gz + ldc * sc[kSsUwall]) / den;
  const double pr = (gr + ldc * 0.0) / den;
  uz[c] = pz;
  ur[c] = pr;
  duz[c] = pz - gz;
  dur[c] = pr - gr;
}

// omega += p, p the interior curl of (du_z, du_r) and 0 on the border.
__global__ void soft_add_curl_kernel(double* __restrict__ w, const double* __restrict__ duz,
                                     const double* __restrict__ dur, int nr, int nz, double dx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  const int64_t c = (int64_t)i * nz + j;
  double p = 0.0;
  if (i >= 1 && i <= nr - 2 && j >= 1 && j <= nz - 2) {
    const double two_dx = 2.0 * dx;
    p = (dur[c + 1] - dur[c - 1]) / two_dx - (duz[c + nz] - duz[c - nz]) / two_dx;
  }
  w[c] = w[c] + p;
}

// eta <- advected eta; phi pinned to the sphere of eta in phi > -3 dx.
__device__ __forceinline__ void soft_pin_at(double* __restrict__ e1, double* __restrict__ e2,
                                            const double* __restrict__ e1n,
                                            const double* __restrict__ e2n,
                                            double* __restrict__ phi, int64_t c, double z_cm,
                                            double r_cm, double r_ball, double dx) {
  const double a = e1n[c], b = e2n[c];
  e1[c] = a;
  e2[c] = b;
  const double da = a - z_cm, db = b - r_cm;
  const double orig = -sqrt(da * da + db * db) + r_ball;
  if (phi[c] > -3 * dx) phi[c] = orig;
}

__global__ void soft_pin_kernel(double* __restrict__ e1, double* __restrict__ e2,
                                const double* __restrict__ e1n, const double* __restrict__ e2n,
                                double* __restrict__ phi, int64_t n, SoftConfig cfg) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  soft_pin_at(e1, e2, e1n, e2n, phi, c, cfg.z_cm, cfg.r_cm, cfg.r_ball, cfg.dx);
}

// chi = smooth Heaviside(phi) (:165-166, written from the upper half of the
// doubled rows), the doubled grid of
// extrapolate_eta_using_least_squares_unb.py:19-24 and the flags of
// extrapolate_using_least_squares.py:33-35 (lower bound 0); block (0, 0)
// also clears the extrapolation sweeps' counters.  Returns whether the cell
// lies in the extrapolation's target zone.
template <typename O>
__device__ __forceinline__ bool soft_lsq_mirror_at(
    const double* __restrict__ phi, O* __restrict__ chi, double moll_zone,
    const double* __restrict__ e1, const double* __restrict__ e2, double* __restrict__ e1d,
    double* __restrict__ e2d, int16_t* __restrict__ cur, int16_t* __restrict__ tgt,
    double* __restrict__ e1d2, double* __restrict__ e2d2, int16_t* __restrict__ cur2, int nr,
    int nz, double extrap_zone, int I, int j) {
  const bool upper = I >= nr;
  const int i = upper ? I - nr : nr - 1 - I;
  const int64_t c = (int64_t)i * nz + j;
  const double p = phi[c];
  const double h = smooth_heaviside(p, moll_zone);
  if (upper) chi[c] = O(h);
  const double in = (h > 0.5) ? 1.0 : 0.0;
  const double pd = -p;
  const int64_t d = (int64_t)I * nz + j;
  const int16_t known = (pd < 0.0) ? 1 : 0;
  const int16_t target = (pd < extrap_zone) ? 1 : 0;
  cur[d] = known;
  tgt[d] = target;
  cur2[d] = known;
  // Values only inside the target zone: a fit weights each value of its
  // window by the cell's known flag and sums from +0, so a cell outside the
  // zone (never known in this step) adds +-0 whatever finite value it holds
  // (the buffers start zeroed and only ever receive finite fits); the copy
  // back gives such cells in * eta (in = 0 there) itself.
  if (!target) return false;
  const double a = in * e1[c];
  const double b = in * e2[c];
  e1d[d] = a;
  e2d[d] = upper ? b : -b;
  // the second buffers of the ping-pong sweeps (cells off the list never
  // change, so both copies must hold them)
  e1d2[d] = a;
  e2d2[d] = upper ? b : -b;
  return true;
}

__global__ void soft_lsq_mirror_kernel(const double* __restrict__ phi,
                                       double* __restrict__ chi, double moll_zone,
                                       const double* __restrict__ e1,
                                       const double* __restrict__ e2, double* __restrict__ e1d,
                                       double* __restrict__ e2d, int16_t* __restrict__ cur,
                                       int16_t* __restrict__ tgt, double* __restrict__ e1d2,
                                       double* __restrict__ e2d2, int16_t* __restrict__ cur2,
                                       int nr, int nz, double extrap_zone,
                                       unsigned* __restrict__ counters, int ncounters) {
  if (blockIdx.x == 0 && blockIdx.y == 0)
    for (int k = threadIdx.x; k < ncounters; k += blockDim.x) counters[k] = 0u;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int I = blockIdx.y;  // doubled row, [0, 2 nr)
  if (j >= nz) return;
  soft_lsq_mirror_at(phi, chi, moll_zone, e1, e2, e1d, e2d, cur, tgt, e1d2, e2d2, cur2, nr, nz,
                     extrap_zone, I, j);
}

// Extrapolated eta back into the fields from whichever ping-pong buffer set
// holds the sweeps' result.
__global__ void soft_lsq_back_kernel(double* __restrict__ e1, double* __restrict__ e2,
                                     const double* __restrict__ e1d,
                                     const double* __restrict__ e2d,
                                     const double* __restrict__ e1d2,
                                     const double* __restrict__ e2d2,
                                     const int16_t* __restrict__ tgt,
                                     const unsigned* __restrict__ counters, int sweeps, int64_t n,
                                     int64_t off) {
  __shared__ int s_second;  // the same for every cell: once per workgroup
  if (threadIdx.x == 0) s_second = lsq_result_is_second(counters, sweeps) ? 1 : 0;
  __syncthreads();
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  if (!tgt[off + c]) {  // outside the target zone: in * eta with in = 0
    e1[c] = 0.0 * e1[c];
    e2[c] = 0.0 * e2[c];
    return;
  }
  const bool second = s_second != 0;
  e1[c] = (second ? e1d2 : e1d)[off + c];
  e2[c] = (second ? e2d2 : e2d)[off + c];
}

__device__ __forceinline__ void soft_finalize_body(double* __restrict__ sc,
                                                   const unsigned* __restrict__ counters,
                                                   int sweeps, unsigned* __restrict__ list_count,
                                                   unsigned* __restrict__ need) {
  *list_count = 0u;  // for the next step's list of cells to extrapolate
  // sweeps that fitted cells: the queued ones, then the tail's (its last
  // sweep fits nothing)
  int used = 0;
  while (used < sweeps && counters[used] != 0u) ++used;
  if (used == sweeps && counters[sweeps] != 0u) {
    used += (int)counters[sweeps] - 1;
    sc[kSsLsqUnconverged] = sc[kSsLsqUnconverged] + 1.0;  // steps the tail finished
  }
  sc[kSsLsqSweeps] = double(used);
  atomicMax(need, (unsigned)used + 1u);  // sweeps this loop took
  sc[kSsT] = sc[kSsT] + sc[kSsDt];
  sc[kSsIt] = sc[kSsIt] + 1.0;
}

__global__ void soft_finalize_kernel(double* __restrict__ sc,
                                     const unsigned* __restrict__ counters, int sweeps,
                                     unsigned* __restrict__ list_count,
                                     unsigned* __restrict__ need, int streaming) {
  if (threadIdx.x != 0) return;
  if (streaming) sc[kSsFreqTimer] = sc[kSsFreqTimer] + sc[kSsDt];
  soft_finalize_body(sc, counters, sweeps, list_count, need);
}

template <typename T>
__global__ void soft_zero_kernel(T* __restrict__ p, int64_t n, int64_t stride) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k * stride] = T(0);
}

// ------------------------------------------------------------------ stepper

// B: storage type of the flow fields (the soft batch's SoftCaseT); the
// single-case stepper is FP64 (B = double).
template <typename B>
struct SoftBuffersT {
  B *vorticity, *psi, *u_z, *u_r;
  double *eta1, *eta2, *phi;
  B* char_func;
  double* scalars;
  const double *r, *z, *kill_wz, *kill_wr;
};
using SoftBuffers = SoftBuffersT<double>;

template <typename B>
struct SoftStepperT {
  SoftConfig cfg{};
  SoftBuffersT<B> buf{};
  StokesPlan* plan = nullptr;
  std::vector<void*> allocs;
  B *w0 = nullptr, *w1 = nullptr, *tmp = nullptr, *spec = nullptr;
  double *e1n = nullptr, *e2n = nullptr, *slots = nullptr;
  double *e1d = nullptr, *e2d = nullptr;
  int16_t *cur = nullptr, *tgt = nullptr;
  B *s11 = nullptr, *s12 = nullptr, *s22 = nullptr;
  B *g1z = nullptr, *g1r = nullptr, *g2z = nullptr, *g2r = nullptr;
  B *tz = nullptr, *tr = nullptr;
  ReinitWork rw{};
  LsqWork lw{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  hipStream_t graph_stream = nullptr;
  // queued-sweep growth: the most sweeps a step has taken (device), its
  // pinned host copy and the event that says the copy has landed
  static constexpr int kMaxQueued = 4096;
  unsigned* need = nullptr;
  unsigned* need_host = nullptr;
  hipEvent_t need_ev = nullptr;
  bool need_pending = false;
  int since_request = 0;  // graph launches queued after the pending copy
  int* zone = nullptr;    // batch cases: the target zone's rows (zone_row)

  template <typename P>
  hipError_t alloc(P** p, size_t bytes) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return e;
    allocs.push_back(q);
    *p = static_cast<P*>(q);
    return hipSuccess;
  }

  // batch_case: the state of one case of a soft batch (any-length row
  // transforms and radial segments of kSoftBatchTriL rows, which every case
  // of a batch launch shares)
  hipError_t init(const SoftConfig& c, const SoftBuffersT<B>& b, bool batch_case = false) {
    cfg = c;
    buf = b;
    plan = new (std::nothrow) StokesPlan();
    if (!plan) return hipErrorOutOfMemory;
    if (batch_case) {
      plan->force_general = true;
      plan->force_L = kSoftBatchTriL;
    }
    // the plan picks the half-length transforms or the any-length ones
    hipError_t e = plan->init(cfg.nr, cfg.nz, cfg.dx);
    if (e == hipSuccess && batch_case && plan->tri_G) e = hipErrorInvalidValue;
    if (e != hipSuccess) return e;
    const size_t n = (size_t)cfg.nr * cfg.nz, nd = 2 * n, D8 = sizeof(double);
    B** flow[] = {&w0, &w1, &tmp, &spec, &s11, &s12, &s22, &g1z, &g1r, &g2z, &g2r, &tz, &tr};
    for (B** f : flow)
      if ((e = alloc(f, n * sizeof(B))) != hipSuccess) return e;
    double** geo[] = {&e1n, &e2n, &rw.d0, &rw.d1, &rw.dist};
    for (double** f : geo)
      if ((e = alloc(f, n * D8)) != hipSuccess) return e;
    if ((e = alloc(&rw.frozen, n)) != hipSuccess) return e;
    if ((e = alloc(&rw.ch0, n)) != hipSuccess) return e;
    if ((e = alloc(&rw.ch1, n)) != hipSuccess) return e;
    const size_t tb = 2 * sizeof(unsigned) * (size_t)reinit_tile_total(cfg.nr, cfg.nz);
    if ((e = alloc(&rw.tiles, tb)) != hipSuccess) return e;
    if ((e = hipMemset(rw.tiles, 0, tb)) != hipSuccess) return e;
    if ((e = alloc(&rw.list, tb)) != hipSuccess) return e;
    if ((e = alloc(&rw.cnt, 3 * sizeof(unsigned))) != hipSuccess) return e;
    if ((e = alloc(&e1d, nd * D8)) != hipSuccess) return e;
    if ((e = alloc(&e2d, nd * D8)) != hipSuccess) return e;
    if ((e = alloc(&cur, nd * sizeof(int16_t))) != hipSuccess) return e;
    if ((e = alloc(&tgt, nd * sizeof(int16_t))) != hipSuccess) return e;
    if ((e = alloc(&lw.counters, sizeof(unsigned) * (kMaxQueued + 1))) != hipSuccess) return e;
    if ((e = alloc(&need, sizeof(unsigned))) != hipSuccess) return e;
    if ((e = hipMemset(need, 0, sizeof(unsigned))) != hipSuccess) return e;
    // every row: the first step a case takes here visits the whole grid
    // (its flags and values are written nowhere else before it), also when
    // its state was loaded from another run
    const int all_rows[4] = {0, cfg.nr - 1, 0, cfg.nr - 1};
    if ((e = alloc(&zone, sizeof(all_rows))) != hipSuccess) return e;
    if ((e = hipMemcpy(zone, all_rows, sizeof(all_rows), hipMemcpyHostToDevice)) != hipSuccess)
      return e;
    if (cfg.lsq_grow) {
      if ((e = hipHostMalloc(reinterpret_cast<void**>(&need_host), sizeof(unsigned),
                             hipHostMallocDefault)) != hipSuccess)
        return e;
      if ((e = hipEventCreateWithFlags(&need_ev, hipEventDisableTiming)) != hipSuccess) return e;
    }
    if ((e = alloc(&lw.list, nd * sizeof(int))) != hipSuccess) return e;
    if ((e = alloc(&lw.tail, 4 * nd * sizeof(int))) != hipSuccess) return e;
    lw.cap = (int)nd;
    if ((e = alloc(&lw.cur2, nd * sizeof(int16_t))) != hipSuccess) return e;
    if ((e = alloc(&lw.ex2, nd * D8)) != hipSuccess) return e;
    if ((e = alloc(&lw.ey2, nd * D8)) != hipSuccess) return e;
    if ((e = alloc(&lw.count, sizeof(unsigned))) != hipSuccess) return e;
    if ((e = hipMemset(lw.count, 0, sizeof(unsigned))) != hipSuccess) return e;
    if ((e = alloc(&slots, D8 * kSlotsPerCase)) != hipSuccess) return e;
    // scratch the reference allocates once and never fully rewrites
    // (gradient and tau scratch keep zeros where they are never written)
    for (B* f : {g1z, g1r, g2z, g2r, tz, tr, s11, s12, s22}) {
      soft_zero_kernel<B><<<ceil_div((int64_t)n, 256), 256>>>(f, (int64_t)n, 1);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    soft_zero_kernel<double><<<1, 64>>>(slots, kMaxSlots, kSlotStride);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // the doubled grid's values: cells outside a step's target zone keep
    // what they hold, which must be finite
    for (double* f : {e1d, e2d, lw.ex2, lw.ey2}) {
      soft_zero_kernel<double><<<ceil_div((int64_t)nd, 256), 256>>>(f, (int64_t)nd, 1);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return hipDeviceSynchronize();
  }

  ~SoftStepperT() {
    drop_graphs();
    if (need_ev) (void)hipEventDestroy(need_ev);
    if (need_host) (void)hipHostFree(need_host);
    for (void* p : allocs) (void)hipFree(p);
    delete plan;
  }

  // The single-case FP64 step (B = double).
  hipError_t enqueue_step(hipStream_t s) {
    static_assert(std::is_same<B, double>::value, "the single-case soft step is FP64");
    const int nr = cfg.nr, nz = cfg.nz;
    const int64_t n = (int64_t)nr * nz;
    const dim3 g2(ceil_div(nz, 256), nr);
    double* sc = buf.scalars;
    hipError_t e;
#define AXS_TRY_STEP(...)                              \
  do {                                                 \
    if ((e = (__VA_ARGS__)) != hipSuccess) return e;   \
  } while (0)
#define AXS_TRY_LAUNCH()                                       \
  do {                                                         \
    if ((e = hipGetLastError()) != hipSuccess) return e;       \
  } while (0)
    AXS_TRY_STEP(launch_taper_dct2_rows<double>(*plan, buf.vorticity, w0, spec, buf.kill_wz,
                                                buf.kill_wr, cfg.kill_width, s));
    AXS_TRY_STEP(launch_tridiag<double>(*plan, spec, s));
    AXS_TRY_STEP(launch_dct3_rows<double>(*plan, spec, buf.psi, nr, s));
    AXS_TRY_STEP(launch_velocity_from_psi<double>(buf.u_z, buf.u_r, buf.psi, buf.r, nr, nz,
                                                  cfg.dx, s));
    const int mb = std::min<int64_t>(ceil_div(n, 256), 1024);
    soft_velocity_max_kernel<<<mb, 256, 0, s>>>(buf.u_z, buf.u_r, n, slots);
    AXS_TRY_LAUNCH();
    soft_dt_kernel<<<1, 64, 0, s>>>(sc, slots, cfg);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(launch_eno3_refmap<double>(buf.eta1, buf.eta2, buf.u_z, buf.u_r, e1n, e2n, nr,
                                            nz, Scalar<double>{0.0, sc + kSsDtByDx}, s));
    soft_pin_kernel<<<ceil_div(n, 256), 256, 0, s>>>(buf.eta1, buf.eta2, e1n, e2n, buf.phi, n,
                                                     cfg);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(launch_reinit(buf.phi, buf.phi, nullptr, nr, nz, cfg.dx, cfg.reinit_band,
                               cfg.reinit_iterations, rw, s));
    if (cfg.streaming)  // soft_sphere_streaming.py:197-198
      AXS_TRY_STEP(launch_eno3_conservative<double>(w0, buf.u_z, buf.u_r, w1, nr, nz,
                                                    Scalar<double>{0.0, sc + kSsDtByDx}, -1.0,
                                                    s));
    soft_lsq_mirror_kernel<<<dim3(ceil_div(nz, 256), 2 * nr), 256, 0, s>>>(
        buf.phi, buf.char_func, cfg.moll_zone, buf.eta1, buf.eta2, e1d, e2d, cur, tgt, lw.ex2,
        lw.ey2, lw.cur2, nr, nz, cfg.extrap_zone, lw.counters, cfg.lsq_sweeps + 1);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(launch_lsq_list(cur, tgt, 2 * n, lw, s));
    // ping-pong sweeps (sweep q reads the first set when q is even); those
    // after the loop has ended return at once
    for (int q = 0; q < cfg.lsq_sweeps; ++q) {
      const bool odd = q & 1;
      AXS_TRY_STEP(launch_lsq_sweep_pingpong(odd ? lw.cur2 : cur, odd ? lw.ex2 : e1d,
                                             odd ? lw.ey2 : e2d, odd ? cur : lw.cur2,
                                             odd ? e1d : lw.ex2, odd ? e2d : lw.ey2, tgt, buf.z,
                                             buf.z, 2 * nr, nz, lw, q, s));
    }
    // a loop longer than the queued sweeps ends in one workgroup
    AXS_TRY_STEP(launch_lsq_sweep_tail(cur, e1d, e2d, tgt, buf.z, buf.z, 2 * nr, nz, lw,
                                       cfg.lsq_sweeps, s));
    soft_lsq_back_kernel<<<ceil_div(n, 256), 256, 0, s>>>(
        buf.eta1, buf.eta2, e1d, e2d, lw.ex2, lw.ey2, tgt, lw.counters, cfg.lsq_sweeps, n, n);
    AXS_TRY_LAUNCH();
    // the TG driver (:186-213) adds the stress to the tapered vorticity and
    // then advects it; the streaming driver (:198-264) advects it first
    // (ENO3 below, w0 -> w1 before the stress) and penalises the core after
    double* wstress = cfg.streaming ? w1 : w0;
    AXS_TRY_STEP(launch_solid_sigma<double>(s11, s12, s22, cfg.G, cfg.dx, buf.eta1, buf.eta2, g1z,
                                            g1r, g2z, g2r, nr, nz, s, buf.char_func));
    AXS_TRY_STEP(launch_div_tau<double>(wstress, tz, tr, s11, s12, s22, buf.r, nr, nz,
                                        Scalar<double>{0.0, sc + kSsDt}, cfg.dx, s));
    if (!cfg.streaming) {
      AXS_TRY_STEP(launch_eno3_conservative<double>(w0, buf.u_z, buf.u_r, w1, nr, nz,
                                                    Scalar<double>{0.0, sc + kSsDtByDx}, -1.0,
                                                    s));
    } else {
      // e1n / e2n (the advected eta, consumed by the pin) hold the
      // correction; the stress scratch keeps its never-written entries
      soft_tether_kernel<<<g2, 256, 0, s>>>(buf.u_z, buf.u_r, e1n, e2n, buf.r, buf.z, sc, cfg);
      AXS_TRY_LAUNCH();
      soft_add_curl_kernel<<<g2, 256, 0, s>>>(w1, e1n, e2n, nr, nz, cfg.dx);
      AXS_TRY_LAUNCH();
    }
    AXS_TRY_STEP(launch_rk2<double, RRow<double>>(
        w1, tmp, buf.vorticity, RRow<double>{buf.r}, nr, nz, Scalar<double>{0.0, sc + kSsHalfNuDt},
        Scalar<double>{0.0, sc + kSsNuDt}, cfg.dx, s));
    soft_finalize_kernel<<<1, 64, 0, s>>>(sc, lw.counters, cfg.lsq_sweeps, lw.count, need,
                                          cfg.streaming);
    AXS_TRY_LAUNCH();
#undef AXS_TRY_STEP
#undef AXS_TRY_LAUNCH
    return hipSuccess;
  }

  // kChunkSteps steps per graph (a graph launch costs microseconds) and a
  // one-step graph for the remainder.
  static constexpr int kChunkSteps = 8;
  hipGraph_t chunk_graph = nullptr;
  hipGraphExec_t chunk_exec = nullptr;

  void drop_graphs() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (chunk_exec) (void)hipGraphExecDestroy(chunk_exec);
    if (chunk_graph) (void)hipGraphDestroy(chunk_graph);
    graph_exec = nullptr;
    graph = nullptr;
    chunk_exec = nullptr;
    chunk_graph = nullptr;
  }

  hipError_t capture(hipStream_t s) {
    drop_graphs();
    for (int q = 0; q < 2; ++q) {
      hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) return e;
      hipError_t le = hipSuccess;
      for (int k = 0; k < (q == 0 ? 1 : kChunkSteps) && le == hipSuccess; ++k)
        le = enqueue_step(s);
      hipGraph_t* g = q == 0 ? &graph : &chunk_graph;
      hipGraphExec_t* x = q == 0 ? &graph_exec : &chunk_exec;
      e = hipStreamEndCapture(s, g);
      if (le != hipSuccess) return le;
      if (e != hipSuccess) return e;
      if ((e = hipGraphInstantiate(x, *g, nullptr, nullptr, 0)) != hipSuccess) return e;
    }
    graph_stream = s;
    return hipSuccess;
  }

  // Queued-sweep growth (cfg.lsq_grow): a graph launch is followed by an
  // asynchronous copy of the most sweeps a step has taken; once it has
  // landed (polled; waited for only when 16 graph launches are queued behind
  // it, so the device keeps ~128 steps of work) the queue grows to that + 4,
  // rounded up to 4, and the graphs are dropped to be captured again.  Loops
  // that outrun the queue meanwhile are finished by the extrapolation tail.
  hipError_t poll_need() {
    if (!need_pending) return hipSuccess;
    hipError_t e = since_request >= 16 ? hipEventSynchronize(need_ev) : hipEventQuery(need_ev);
    if (e == hipErrorNotReady) return hipSuccess;
    if (e != hipSuccess) return e;
    need_pending = false;
    const int want = std::min(kMaxQueued, (int)((*need_host + 4u + 3u) / 4u * 4u));
    if (want > cfg.lsq_sweeps) {
      cfg.lsq_sweeps = want;
      drop_graphs();
    }
    return hipSuccess;
  }
  hipError_t request_need(hipStream_t s) {
    if (need_pending) return hipSuccess;
    hipError_t e = hipMemcpyAsync(need_host, need, sizeof(unsigned), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipEventRecord(need_ev, s);
    if (e == hipSuccess) {
      need_pending = true;
      since_request = 0;
    }
    return e;
  }

  hipError_t run(int nsteps, bool use_graph, hipStream_t s) {
    if (!use_graph) {
      for (int q = 0; q < nsteps; ++q) {
        hipError_t e = enqueue_step(s);
        if (e != hipSuccess) return e;
      }
      return hipSuccess;
    }
    if (graph_stream != s) drop_graphs();
    hipError_t e = hipSuccess;
    int q = 0;
    while (q < nsteps) {
      if (cfg.lsq_grow && (e = poll_need()) != hipSuccess) return e;
      if (!graph_exec && (e = capture(s)) != hipSuccess) return e;
      const bool chunk = q + kChunkSteps <= nsteps;
      if ((e = hipGraphLaunch(chunk ? chunk_exec : graph_exec, s)) != hipSuccess) return e;
      q += chunk ? kChunkSteps : 1;
      ++since_request;
      if (cfg.lsq_grow && (e = request_need(s)) != hipSuccess) return e;
    }
    return hipSuccess;
  }
};
using SoftStepper = SoftStepperT<double>;

// ------------------------------------------------------------------ batches
// Kernels over every case of a soft batch: one row segment of a case per
// workgroup (the batch's block tables), or one block per case.

template <typename B>
__global__ void soft_batch_dt_kernel(const SoftCaseT<B>* __restrict__ cs) {
  const SoftCaseT<B>& C = cs[blockIdx.x];
  if (threadIdx.x == 0) {  // the last step's zone kept, an empty one to record
    C.zone[2] = C.zone[0];
    C.zone[3] = C.zone[1];
    C.zone[0] = C.nr;
    C.zone[1] = -1;
  }
  soft_dt_body(C.scalars, C.slots, C.CFL, C.dx, C.eps, C.G, C.rho_f, C.nu);
}

template <typename B>
__global__ void soft_batch_pin_kernel(const SoftCaseT<B>* __restrict__ cs,
                                      const uint32_t* __restrict__ blocks) {
  const SoftBlock b = soft_block(blocks);
  const SoftCaseT<B>& C = cs[b.c];
  if (b.j >= C.nz || !prev_zone_row(C, b.i)) return;
  soft_pin_at(C.eta1, C.eta2, C.e1n, C.e2n, C.phi, (int64_t)b.i * C.nz + b.j, C.z_cm, C.r_cm,
              C.r_ball, C.dx);
}

template <typename B>
__global__ void soft_batch_mirror_kernel(const SoftCaseT<B>* __restrict__ cs,
                                         const uint32_t* __restrict__ blocks, int ncounters) {
  const SoftBlock b = soft_block(blocks);  // a row of the doubled grid
  const SoftCaseT<B>& C = cs[b.c];
  if (b.i == 0 && b.j < 256)  // the case's first workgroup clears its sweep counters
    for (int k = threadIdx.x; k < ncounters; k += blockDim.x) C.lw.counters[k] = 0u;
  // rows beyond the last zone keep their flags (0) and chi (0) from the
  // step that last visited them (the whole workgroup leaves together)
  if (!prev_zone_row(C, b.i >= C.nr ? b.i - C.nr : C.nr - 1 - b.i)) return;
  bool target = false;
  if (b.j < C.nz)
    target = soft_lsq_mirror_at(C.phi, C.chi, C.moll_zone, C.eta1, C.eta2, C.e1d, C.e2d, C.cur,
                                C.tgt, C.lw.ex2, C.lw.ey2, C.lw.cur2, C.nr, C.nz,
                                C.extrap_zone, b.i, b.j);
  // the workgroup is one row segment: its row joins the zone if any cell
  // of the segment is a target (upper half: the physical rows)
  if (__syncthreads_or(target) && b.i >= C.nr && threadIdx.x == 0) {
    atomicMin(C.zone, b.i - C.nr);
    atomicMax(C.zone + 1, b.i - C.nr);
  }
}

template <typename B>
__global__ void soft_batch_back_kernel(const SoftCaseT<B>* __restrict__ cs,
                                       const uint32_t* __restrict__ blocks, int sweeps) {
  __shared__ int s_second;
  const SoftBlock b = soft_block(blocks);
  const SoftCaseT<B>& C = cs[b.c];
  if (!zone_row(C, b.i)) return;  // the whole workgroup (one row)
  if (threadIdx.x == 0) s_second = lsq_result_is_second(C.lw.counters, sweeps) ? 1 : 0;
  __syncthreads();
  if (b.j >= C.nz) return;
  const int64_t n = (int64_t)C.nr * C.nz, c = (int64_t)b.i * C.nz + b.j;
  if (!C.tgt[n + c]) {  // outside the target zone: in * eta with in = 0
    C.eta1[c] = 0.0 * C.eta1[c];
    C.eta2[c] = 0.0 * C.eta2[c];
    return;
  }
  const bool second = s_second != 0;
  C.eta1[c] = (second ? C.lw.ex2 : C.e1d)[n + c];
  C.eta2[c] = (second ? C.lw.ey2 : C.e2d)[n + c];
}

template <typename B>
__global__ void soft_batch_finalize_kernel(const SoftCaseT<B>* __restrict__ cs, int sweeps) {
  if (threadIdx.x != 0) return;
  const SoftCaseT<B>& C = cs[blockIdx.x];
  soft_finalize_body(C.scalars, C.lw.counters, sweeps, C.lw.count, C.need);
}

// Many soft cases stepped together: each pass of the step is one launch over
// all cases (SoftCaseT table on the device), in the single-case step's
// order, so a case's results are the same bits alone and in any mix.  Cases
// join and leave between runs; the step graphs are captured again after a
// change.  The extrapolation queue (sweeps launched per step) is the
// batch's, grown from the most sweeps any case has taken as in the single
// stepper; the result never depends on it (a longer loop is finished by each
// case's tail workgroup).
template <typename B>
struct SoftBatchT {
  struct Entry {
    int id;
    SoftStepperT<B>* st;
  };
  std::vector<Entry> cases;
  int next_id = 1;
  int max_cases = 0;
  SoftCaseT<B>* d_cases = nullptr;
  std::vector<SoftCaseT<B>> h_cases;
  SoftBatchDims dims{};
  int sweeps = 12;
  unsigned* need = nullptr;
  unsigned* need_host = nullptr;
  hipEvent_t need_ev = nullptr;
  bool need_pending = false;
  int since_request = 0;
  static constexpr int kChunkSteps = 8;
  hipGraph_t graph = nullptr, chunk_graph = nullptr;
  hipGraphExec_t graph_exec = nullptr, chunk_exec = nullptr;
  hipStream_t graph_stream = nullptr;
  bool table_dirty = true;
  int64_t chunk_launches = 0, step_launches = 0, captures = 0;
  uint32_t* d_blocks = nullptr;  // cell then doubled-grid block tables
  size_t blocks_cap = 0;
  unsigned* d_rlist = nullptr;  // reinit work lists (two of the tile table's size)
  unsigned* d_rcnt = nullptr;   // their three counters
  size_t rlist_cap = 0;

  hipError_t init(int cap) {
    max_cases = cap;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&d_cases), sizeof(SoftCaseT<B>) * cap);
    if (e != hipSuccess) return e;
    if ((e = hipMalloc(reinterpret_cast<void**>(&need), sizeof(unsigned))) != hipSuccess) return e;
    if ((e = hipMemset(need, 0, sizeof(unsigned))) != hipSuccess) return e;
    if ((e = hipMalloc(reinterpret_cast<void**>(&d_rcnt), 3 * sizeof(unsigned))) != hipSuccess)
      return e;
    if ((e = hipHostMalloc(reinterpret_cast<void**>(&need_host), sizeof(unsigned),
                           hipHostMallocDefault)) != hipSuccess)
      return e;
    return hipEventCreateWithFlags(&need_ev, hipEventDisableTiming);
  }

  void drop_graphs() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (chunk_exec) (void)hipGraphExecDestroy(chunk_exec);
    if (chunk_graph) (void)hipGraphDestroy(chunk_graph);
    graph_exec = nullptr;
    graph = nullptr;
    chunk_exec = nullptr;
    chunk_graph = nullptr;
  }

  ~SoftBatchT() {
    (void)hipDeviceSynchronize();
    drop_graphs();
    for (Entry& c : cases) delete c.st;
    if (need_ev) (void)hipEventDestroy(need_ev);
    if (need_host) (void)hipHostFree(need_host);
    if (need) (void)hipFree(need);
    if (d_cases) (void)hipFree(d_cases);
    if (d_blocks) (void)hipFree(d_blocks);
    if (d_rlist) (void)hipFree(d_rlist);
    if (d_rcnt) (void)hipFree(d_rcnt);
  }

  static SoftCaseT<B> describe(const SoftStepperT<B>& S, unsigned* need) {
    SoftCaseT<B> C{};
    const SoftConfig& g = S.cfg;
    C.nr = g.nr;
    C.nz = g.nz;
    C.kill_width = g.kill_width;
    C.reinit_iterations = g.reinit_iterations;
    C.dx = g.dx;
    C.nu = g.nu;
    C.G = g.G;
    C.rho_f = g.rho_f;
    C.CFL = g.CFL;
    C.eps = g.eps;
    C.r_ball = g.r_ball;
    C.z_cm = g.z_cm;
    C.r_cm = g.r_cm;
    C.moll_zone = g.moll_zone;
    C.extrap_zone = g.extrap_zone;
    C.reinit_band = g.reinit_band;
    const SoftBuffersT<B>& b = S.buf;
    C.vorticity = b.vorticity;
    C.psi = b.psi;
    C.u_z = b.u_z;
    C.u_r = b.u_r;
    C.eta1 = b.eta1;
    C.eta2 = b.eta2;
    C.phi = b.phi;
    C.chi = b.char_func;
    C.scalars = b.scalars;
    C.r = b.r;
    C.z = b.z;
    C.kill_wz = b.kill_wz;
    C.kill_wr = b.kill_wr;
    C.w0 = S.w0;
    C.w1 = S.w1;
    C.tmp = S.tmp;
    C.spec = S.spec;
    C.e1n = S.e1n;
    C.e2n = S.e2n;
    C.slots = S.slots;
    C.e1d = S.e1d;
    C.e2d = S.e2d;
    C.cur = S.cur;
    C.tgt = S.tgt;
    C.s11 = S.s11;
    C.s12 = S.s12;
    C.s22 = S.s22;
    C.g1z = S.g1z;
    C.g1r = S.g1r;
    C.g2z = S.g2z;
    C.g2r = S.g2r;
    C.tz = S.tz;
    C.tr = S.tr;
    C.rw = S.rw;
    C.lw = S.lw;
    C.need = need;
    C.zone = S.zone;
    const StokesPlan& P = *S.plan;
    C.any = P.any->args;
    C.ge2 = P.ge2;
    C.te1 = P.e1;
    C.te2 = P.e2;
    C.any_half = P.any_half ? 1 : 0;
    C.ta = P.a;
    C.tb = P.b;
    C.tc = P.c;
    C.tlam = P.lam;
    C.tcseg = P.cseg;
    C.rvec = P.rvec;
    C.tK = P.tri_K;
    C.tP = P.tri_P;
    return C;
  }

  // Device table and launch shapes of the current cases.
  hipError_t refresh() {
    h_cases.clear();
    dims = SoftBatchDims{};
    for (const Entry& c : cases) {
      const SoftStepperT<B>& S = *c.st;
      h_cases.push_back(describe(S, need));
      const StokesPlan& P = *S.plan;
      dims.max_nr = std::max(dims.max_nr, P.nr);
      dims.max_nz = std::max(dims.max_nz, P.nz);
      dims.max_chunks = std::max(dims.max_chunks, (int)ceil_div(P.nz, P.tri_K));
      // a Stockham stage gives every butterfly its own lane slot in one
      // pass (at most 4 per thread, no loop): the widest stage, L / (the
      // smallest radix) butterflies, sets the row workgroup, ~4 per thread
      const int wide = widest_stage(P.any->args);
      dims.any_nth =
          std::max(dims.any_nth, std::min(1024, (int)(ceil_div(ceil_div(wide, 4), 64) * 64)));
      dims.any_lds = std::max(dims.any_lds, P.any->lds_bytes<B>());
      dims.tri_lds =
          std::max(dims.tri_lds, tri_lds_bytes(P.nr, kSoftBatchTriL, P.tri_K, P.tri_P, false));
    }
    for (const Entry& c : cases) {
      const int per = (int)ceil_div(widest_stage(c.st->plan->any->args), dims.any_nth);
      if (per > 4) return hipErrorInvalidValue;  // a stage would leave butterflies out
      dims.any_per = std::max(dims.any_per, per <= 1 ? 1 : (per <= 2 ? 2 : 4));
    }
    dims.ncase = (int)cases.size();
    // block tables: every 256-cell row segment of every case, over its grid
    // and over its doubled grid
    std::vector<uint32_t> tab;
    for (int c = 0; c < dims.ncase; ++c) {
      const SoftCaseT<B>& C = h_cases[c];
      for (int i = 0; i < C.nr; ++i)
        for (int b = 0; b < (int)ceil_div(C.nz, 256); ++b) tab.push_back(soft_block_entry(c, i, b));
    }
    dims.ncell_blocks = (int)tab.size();
    for (int c = 0; c < dims.ncase; ++c) {
      const SoftCaseT<B>& C = h_cases[c];
      for (int i = 0; i < 2 * C.nr; ++i)
        for (int b = 0; b < (int)ceil_div(C.nz, 256); ++b) tab.push_back(soft_block_entry(c, i, b));
    }
    dims.ndbl_blocks = (int)tab.size() - dims.ncell_blocks;
    const size_t tiles0 = tab.size();
    for (int c = 0; c < dims.ncase; ++c) {
      SoftCaseT<B>& C = h_cases[c];
      C.reinit_tile_base = (int)(tab.size() - tiles0);
      for (int ti = 0; ti < (int)ceil_div(C.nr, kReinitTileRows); ++ti)
        for (int tj = 0; tj < (int)ceil_div(C.nz, kReinitTileCols); ++tj)
          tab.push_back(soft_block_entry(c, ti, tj));
    }
    dims.nreinit_tiles = (int)(tab.size() - tiles0);
    drop_graphs();
    table_dirty = false;
    if (cases.empty()) return hipSuccess;
    hipError_t e;
    if (tab.size() > blocks_cap) {
      if (d_blocks) (void)hipFree(d_blocks);
      d_blocks = nullptr;
      blocks_cap = 0;
      if ((e = hipMalloc(reinterpret_cast<void**>(&d_blocks), sizeof(uint32_t) * tab.size())) !=
          hipSuccess)
        return e;
      blocks_cap = tab.size();
    }
    if ((e = hipMemcpy(d_blocks, tab.data(), sizeof(uint32_t) * tab.size(),
                       hipMemcpyHostToDevice)) != hipSuccess)
      return e;
    if ((size_t)dims.nreinit_tiles > rlist_cap) {
      if (d_rlist) (void)hipFree(d_rlist);
      d_rlist = nullptr;
      rlist_cap = 0;
      if ((e = hipMalloc(reinterpret_cast<void**>(&d_rlist),
                         2 * sizeof(unsigned) * (size_t)dims.nreinit_tiles)) != hipSuccess)
        return e;
      rlist_cap = (size_t)dims.nreinit_tiles;
    }
    dims.reinit_list = d_rlist;
    dims.reinit_cnt = d_rcnt;
    dims.cell_blocks = d_blocks;
    dims.dbl_blocks = d_blocks + dims.ncell_blocks;
    dims.reinit_tiles = d_blocks + tiles0;
    return hipMemcpy(d_cases, h_cases.data(), sizeof(SoftCaseT<B>) * cases.size(),
                     hipMemcpyHostToDevice);
  }

  hipError_t add(const SoftConfig& c, const SoftBuffersT<B>& b, int* id) {
    if ((int)cases.size() >= max_cases) return hipErrorInvalidValue;
    SoftStepperT<B>* st = new (std::nothrow) SoftStepperT<B>();
    if (!st) return hipErrorOutOfMemory;
    SoftConfig cc = c;
    cc.lsq_sweeps = kMaxQueued;  // counters sized for the largest queue
    cc.lsq_grow = 0;
    hipError_t e = st->init(cc, b, true);
    if (e != hipSuccess) {
      delete st;
      return e;
    }
    // a queued step of the old set may still read the table
    if ((e = hipDeviceSynchronize()) != hipSuccess) {
      delete st;
      return e;
    }
    cases.push_back(Entry{next_id, st});
    *id = next_id++;
    table_dirty = true;
    return hipSuccess;
  }

  hipError_t remove(int id) {
    for (size_t k = 0; k < cases.size(); ++k) {
      if (cases[k].id != id) continue;
      // the case's scratch may still be in use by queued steps
      hipError_t e = hipDeviceSynchronize();
      if (e != hipSuccess) return e;
      delete cases[k].st;
      cases.erase(cases.begin() + k);
      table_dirty = true;
      return hipSuccess;
    }
    return hipErrorInvalidValue;
  }

  static constexpr int kMaxQueued = SoftStepper::kMaxQueued;

  hipError_t enqueue_step(hipStream_t s) {
    const SoftCaseT<B>* T = d_cases;
    const SoftBatchDims& d = dims;
    hipError_t e;
#define AXS_TRY_STEP(...)                            \
  do {                                               \
    if ((e = (__VA_ARGS__)) != hipSuccess) return e; \
  } while (0)
#define AXS_TRY_LAUNCH()                                 \
  do {                                                   \
    if ((e = hipGetLastError()) != hipSuccess) return e; \
  } while (0)
    int iters = 0;
    for (const SoftCaseT<B>& c : h_cases) iters = std::max(iters, c.reinit_iterations);
    AXS_TRY_STEP(launch_soft_batch_dct2(T, d, s));
    AXS_TRY_STEP(launch_soft_batch_tridiag(T, d, s));
    AXS_TRY_STEP(launch_soft_batch_dct3(T, d, s));
    AXS_TRY_STEP(launch_soft_batch_velocity(T, d, s));
    soft_batch_dt_kernel<B><<<d.ncase, 64, 0, s>>>(T);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(launch_soft_batch_refmap(T, d, s));
    soft_batch_pin_kernel<B><<<d.ncell_blocks, 256, 0, s>>>(T, d.cell_blocks);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(launch_soft_batch_reinit(T, d, iters, s));
    soft_batch_mirror_kernel<B><<<d.ndbl_blocks, 256, 0, s>>>(T, d.dbl_blocks, sweeps + 1);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(launch_soft_batch_lsq_list(T, d, s));
    for (int q = 0; q < sweeps; ++q) AXS_TRY_STEP(launch_soft_batch_lsq_sweep(T, d, q, s));
    AXS_TRY_STEP(launch_soft_batch_lsq_tail(T, d, sweeps, s));
    soft_batch_back_kernel<B><<<d.ncell_blocks, 256, 0, s>>>(T, d.cell_blocks, sweeps);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(launch_soft_batch_stress(T, d, s));
    AXS_TRY_STEP(launch_soft_batch_advect(T, d, s));
    AXS_TRY_STEP(launch_soft_batch_diffuse(T, d, s));
    soft_batch_finalize_kernel<B><<<d.ncase, 64, 0, s>>>(T, sweeps);
    AXS_TRY_LAUNCH();
#undef AXS_TRY_STEP
#undef AXS_TRY_LAUNCH
    return hipSuccess;
  }

  hipError_t capture(hipStream_t s) {
    drop_graphs();
    ++captures;
    for (int q = 0; q < 2; ++q) {
      hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) return e;
      hipError_t le = hipSuccess;
      for (int k = 0; k < (q == 0 ? 1 : kChunkSteps) && le == hipSuccess; ++k)
        le = enqueue_step(s);
      hipGraph_t* gr = q == 0 ? &graph : &chunk_graph;
      hipGraphExec_t* x = q == 0 ? &graph_exec : &chunk_exec;
      e = hipStreamEndCapture(s, gr);
      if (le != hipSuccess) return le;
      if (e != hipSuccess) return e;
      if ((e = hipGraphInstantiate(x, *gr, nullptr, nullptr, 0)) != hipSuccess) return e;
    }
    graph_stream = s;
    return hipSuccess;
  }

  // Queue growth as in SoftStepper::poll_need.
  hipError_t poll_need() {
    if (!need_pending) return hipSuccess;
    hipError_t e = since_request >= 16 ? hipEventSynchronize(need_ev) : hipEventQuery(need_ev);
    if (e == hipErrorNotReady) return hipSuccess;
    if (e != hipSuccess) return e;
    need_pending = false;
    const int want = std::min(kMaxQueued, (int)((*need_host + 4u + 3u) / 4u * 4u));
    if (want > sweeps) {
      sweeps = want;
      drop_graphs();
    }
    return hipSuccess;
  }
  hipError_t request_need(hipStream_t s) {
    if (need_pending) return hipSuccess;
    hipError_t e = hipMemcpyAsync(need_host, need, sizeof(unsigned), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipEventRecord(need_ev, s);
    if (e == hipSuccess) {
      need_pending = true;
      since_request = 0;
    }
    return e;
  }

  hipError_t run(int nsteps, bool use_graph, hipStream_t s) {
    hipError_t e;
    if (table_dirty && (e = refresh()) != hipSuccess) return e;
    if (cases.empty() || nsteps == 0) return hipSuccess;
    if (!use_graph) {
      for (int q = 0; q < nsteps; ++q)
        if ((e = enqueue_step(s)) != hipSuccess) return e;
      return hipSuccess;
    }
    if (graph_stream != s) drop_graphs();
    int q = 0;
    while (q < nsteps) {
      if ((e = poll_need()) != hipSuccess) return e;
      if (!graph_exec && (e = capture(s)) != hipSuccess) return e;
      const bool chunk = q + kChunkSteps <= nsteps;
      if ((e = hipGraphLaunch(chunk ? chunk_exec : graph_exec, s)) != hipSuccess) return e;
      ++(chunk ? chunk_launches : step_launches);
      q += chunk ? kChunkSteps : 1;
      ++since_request;
      if ((e = request_need(s)) != hipSuccess) return e;
    }
    return hipSuccess;
  }
};

}  // namespace axs

using namespace axs;

// A soft batch of one storage type: FP64 (axs_soft_batch_create) or FP32
// bulk (axs_soft_batch_create_f32); the other member stays empty.
struct axs_soft_batch {
  bool f32 = false;
  SoftBatchT<double> d;
  SoftBatchT<float> f;
  template <typename F>
  auto with(F&& fn) {
    return f32 ? fn(f) : fn(d);
  }
  template <typename F>
  auto with(F&& fn) const {
    return f32 ? fn(f) : fn(d);
  }
};

struct axs_soft_stepper {
  SoftStepper impl;
};

// Checked copy of a soft configuration and its buffers (axs_soft_buffers or
// axs_soft_buffers_f32).
template <class In, typename B>
static axs_status soft_config_from(const axs_soft_config* c, const In* b, SoftConfig& cfg,
                                   SoftBuffersT<B>& bb) {
  AXS_CHECK_ARG(c && b, "null pointer");
  AXS_CHECK_ARG(c->nr >= 4 && c->nz >= 8 && c->nz % 2 == 0, "need nr >= 4 and even nz >= 8");
  AXS_CHECK_ARG(c->nr <= 4096 && stokes_nz_supported(c->nz),
                "grid too large for the row/column passes");
  AXS_CHECK_ARG(c->nz == 2 * c->nr,
                "the least-squares extrapolation uses z as both grid axes: need nz = 2 nr");
  AXS_CHECK_ARG(c->kill_width >= 2 && 2 * c->kill_width <= c->nz && c->kill_width < c->nr,
                "bad kill_width for the grid");
  AXS_CHECK_ARG(c->lsq_sweeps >= 1 && c->lsq_sweeps <= 4096, "lsq_sweeps must be in [1, 4096]");
  AXS_CHECK_ARG(c->dx > 0 && c->nu > 0 && c->G > 0 && c->rho_f > 0 && c->CFL > 0,
                "dx, nu, G, rho_f and CFL must be positive");
  AXS_CHECK_ARG(c->moll_zone > 0 && c->extrap_zone > 0 && c->reinit_band > 0,
                "moll_zone, extrap_zone and reinit_band must be positive");
  AXS_CHECK_ARG(b->vorticity && b->psi && b->u_z && b->u_r && b->eta1 && b->eta2 && b->phi &&
                    b->char_func && b->scalars && b->r && b->z && b->kill_wz && b->kill_wr,
                "null buffer");
  cfg = SoftConfig{};
  cfg.nr = c->nr;
  cfg.nz = c->nz;
  cfg.kill_width = c->kill_width;
  cfg.lsq_sweeps = c->lsq_sweeps;
  cfg.reinit_iterations = c->reinit_iterations;
  cfg.lsq_grow = c->lsq_grow;
  cfg.dx = c->dx;
  cfg.nu = c->nu;
  cfg.G = c->G;
  cfg.rho_f = c->rho_f;
  cfg.CFL = c->CFL;
  cfg.eps = c->eps;
  cfg.r_ball = c->r_ball;
  cfg.z_cm = c->z_cm;
  cfg.r_cm = c->r_cm;
  cfg.moll_zone = c->moll_zone;
  cfg.extrap_zone = c->extrap_zone;
  cfg.reinit_band = c->reinit_band;
  cfg.streaming = c->streaming != 0;
  if (cfg.streaming) {
    AXS_CHECK_ARG(c->brink_lam >= 0 && c->tether_rad > 0 && c->freq_limit > 0 && c->t_end > 0,
                  "streaming needs brink_lam >= 0 and positive tether_rad, freq_limit, t_end");
    cfg.brink_lam = c->brink_lam;
    cfg.U0 = c->U0;
    cfg.omega = c->omega;
    cfg.e_amp = c->e_amp;
    cfg.tether_rad = c->tether_rad;
    cfg.freq_limit = c->freq_limit;
    cfg.t_end = c->t_end;
  }
  if (cfg.reinit_iterations < 0) cfg.reinit_iterations = reinit_iterations(cfg.reinit_band, cfg.dx);
  AXS_CHECK_ARG(cfg.reinit_iterations <= 1 << 16, "too many reinit iterations");
  bb = SoftBuffersT<B>{b->vorticity, b->psi,     b->u_z, b->u_r, b->eta1,    b->eta2,   b->phi,
                       b->char_func, b->scalars, b->r,   b->z,   b->kill_wz, b->kill_wr};
  return AXS_OK;
}

static axs_status soft_batch_create(int max_cases, bool f32, axs_soft_batch** out) {
  AXS_CHECK_ARG(out, "null pointer");
  *out = nullptr;
  AXS_CHECK_ARG(max_cases >= 1 && max_cases <= kSoftBatchMaxCases,
                "max_cases must be in [1, 1024]");
  axs_soft_batch* bt = new (std::nothrow) axs_soft_batch();
  if (!bt) {
    set_error("out of host memory");
    return AXS_ERR_ALLOC;
  }
  bt->f32 = f32;
  hipError_t e = bt->with([&](auto& b) { return b.init(max_cases); });
  if (e != hipSuccess) {
    set_error(std::string("soft batch: ") + hipGetErrorString(e));
    delete bt;
    return AXS_ERR_HIP;
  }
  *out = bt;
  return AXS_OK;
}

template <typename B, class In>
static axs_status soft_batch_add(SoftBatchT<B>& bt, const axs_soft_config* c, const In* b,
                                 int* case_id) {
  AXS_CHECK_ARG(case_id, "null pointer");
  AXS_CHECK_ARG((int)bt.cases.size() < bt.max_cases, "the batch is full");
  SoftConfig cfg;
  SoftBuffersT<B> bb;
  const axs_status st0 = soft_config_from(c, b, cfg, bb);
  if (st0 != AXS_OK) return st0;
  AXS_CHECK_ARG(!cfg.streaming, "the soft batch steps the Taylor-Green loop (streaming = 0)");
  AXS_CHECK_ARG(any_dft_length(cfg.nz, nullptr) <= kAnyMaxL && cfg.nr <= 1024 * kSoftBatchTriL,
                "grid too large for a soft batch");
  hipError_t e = bt.add(cfg, bb, case_id);
  if (e != hipSuccess) {
    set_error(std::string("soft batch add: ") + hipGetErrorString(e));
    return AXS_ERR_HIP;
  }
  return AXS_OK;
}

extern "C" {

axs_status axs_soft_stepper_create(const axs_soft_config* c, const axs_soft_buffers* b,
                                   axs_soft_stepper** out) {
  AXS_CHECK_ARG(out, "null pointer");
  *out = nullptr;
  SoftConfig cfg;
  SoftBuffers bb;
  const axs_status st0 = soft_config_from(c, b, cfg, bb);
  if (st0 != AXS_OK) return st0;
  axs_soft_stepper* st = new (std::nothrow) axs_soft_stepper();
  if (!st) {
    set_error("out of host memory");
    return AXS_ERR_ALLOC;
  }
  hipError_t e = st->impl.init(cfg, bb);
  if (e != hipSuccess) {
    set_error(std::string("soft stepper: ") + hipGetErrorString(e));
    delete st;
    return AXS_ERR_HIP;
  }
  *out = st;
  return AXS_OK;
}

axs_status axs_soft_batch_create(int max_cases, axs_soft_batch** out) {
  return soft_batch_create(max_cases, false, out);
}

axs_status axs_soft_batch_create_f32(int max_cases, axs_soft_batch** out) {
  return soft_batch_create(max_cases, true, out);
}

axs_status axs_soft_batch_add(axs_soft_batch* bt, const axs_soft_config* c,
                              const axs_soft_buffers* b, int* case_id) {
  AXS_CHECK_ARG(bt, "null batch");
  AXS_CHECK_ARG(!bt->f32, "an FP32-bulk batch takes axs_soft_batch_add_f32");
  return soft_batch_add(bt->d, c, b, case_id);
}

axs_status axs_soft_batch_add_f32(axs_soft_batch* bt, const axs_soft_config* c,
                                  const axs_soft_buffers_f32* b, int* case_id) {
  AXS_CHECK_ARG(bt, "null batch");
  AXS_CHECK_ARG(bt->f32, "an FP64 batch takes axs_soft_batch_add");
  return soft_batch_add(bt->f, c, b, case_id);
}

axs_status axs_soft_batch_remove(axs_soft_batch* bt, int case_id) {
  AXS_CHECK_ARG(bt, "null batch");
  const bool known = bt->with([&](const auto& b) {
    for (const auto& c : b.cases)
      if (c.id == case_id) return true;
    return false;
  });
  AXS_CHECK_ARG(known, "no such case in the batch");
  AXS_HIP(bt->with([&](auto& b) { return b.remove(case_id); }));
  return AXS_OK;
}

axs_status axs_soft_batch_run(axs_soft_batch* bt, int nsteps, int use_graph, void* stream) {
  AXS_CHECK_ARG(bt, "null batch");
  AXS_CHECK_ARG(nsteps >= 0, "negative step count");
  AXS_CHECK_ARG(!use_graph || stream, "graph capture needs a non-null stream");
  AXS_HIP(bt->with([&](auto& b) { return b.run(nsteps, use_graph != 0, as_stream(stream)); }));
  return AXS_OK;
}

axs_status axs_soft_batch_info(const axs_soft_batch* bt, int* ncase, int* queued_sweeps,
                               int64_t* graph_launches, int64_t* captures) {
  AXS_CHECK_ARG(bt, "null batch");
  bt->with([&](const auto& b) {
    if (ncase) *ncase = (int)b.cases.size();
    if (queued_sweeps) *queued_sweeps = b.sweeps;
    if (graph_launches) *graph_launches = b.chunk_launches + b.step_launches;
    if (captures) *captures = b.captures;
    return 0;
  });
  return AXS_OK;
}

axs_status axs_soft_batch_destroy(axs_soft_batch* bt) {
  delete bt;
  return AXS_OK;
}

axs_status axs_soft_stepper_destroy(axs_soft_stepper* st) {
  delete st;
  return AXS_OK;
}

axs_status axs_soft_stepper_run(axs_soft_stepper* st, int nsteps, int use_graph, void* stream) {
  AXS_CHECK_ARG(st, "null stepper");
  AXS_CHECK_ARG(nsteps >= 0, "negative step count");
  AXS_CHECK_ARG(!use_graph || stream, "graph capture needs a non-null stream");
  AXS_HIP(st->impl.run(nsteps, use_graph != 0, as_stream(stream)));
  return AXS_OK;
}

}  // extern "C"
