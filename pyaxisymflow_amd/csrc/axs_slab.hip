// Device-resident timestep of the periodic rigid-slab driver
// (examples/PeriodicRigidSlab/periodic_rigid_slab.py:93-158): a rigid annular
// wall oscillating along z (Brinkman penalisation towards U_0 cos(omega t))
// in a z-periodic domain whose columns 0..g-1 and nz-g..nz-1 are ghost copies
// of the periodic interior.  One step, in the reference's order:
//   kill_r        kill taper of the top kill_width rows and row 0 of omega
//                 (kill_boundary_vorticity_sine.py:17-27; weights from the
//                 host)
//   gather        omega's interior columns into the solve's right-hand side
//   rfft rows, tridiag, irfft rows
//                 psi_inner: the periodic-z Stokes solve
//                 (FastDiagonalisationStokesSolver.py:88-97, Neumann r)
//   scatter       psi's interior columns
//   ghost(psi), velocity, velocity max
//                 compute_velocity_from_psi_periodic (:20-35)
//   dt            min(0.9 dx^2 / 4 / nu, CFL dx / (max + eps)), U = U_0 cos(omega t)
//   penalise      u_upen = u; u = Brinkman(u_upen) (the penalised velocity
//                 is what the step advects with); du = u - u_upen
//   ghost(du_r, du_z), curl
//                 omega += the interior curl of du (compute_vorticity_from_
//                 velocity_periodic, :16-28)
//   ghost(u_z, u_r, omega), ENO3
//                 advect_vorticity_via_eno3_periodic (:47-91) -> w1
//   ghost(w1), RK2 stage 1, ghost(tmp), RK2 stage 2
//                 diffusion_RK2_periodic (:47-90) -> omega
//   finalize      t += dt, it += 1
// FP64; every scalar on the device; steps captured into hipGraphs (no memset
// nodes).
#include <algorithm>
#include <cmath>
#include <limits>
#include <new>
#include <vector>

#include "axs_eno3.hpp"
#include "axs_fused.hpp"
#include "axs_stencil.hpp"
#include "axs_stokes.hpp"

namespace axs {

enum : int {
  kSlT = 0,     // time
  kSlDt,        // dt of the current step
  kSlDtByDx,    // dt / dx
  kSlHalfNuDt,  // 0.5 nu dt
  kSlNuDt,      // nu dt
  kSlU,         // wall velocity U_0 cos(omega t)
  kSlMaxU,      // max(|u_z| + |u_r|)
  kSlIt,        // step counter
  kSlCount
};
static_assert(kSlCount <= AXS_RIGID_SCALARS, "slab scalar slots exceed the array");

struct SlabConfig {
  int nr, nz, ghost, kill_width;
  double dx, nu, lam, U0, omega, CFL, eps;
};

struct SlabBuffers {
  double *vorticity, *psi, *u_z, *u_r, *scalars;
  const double *char_func, *r, *kill_wr;
};

// ------------------------------------------------------------------ kernels

// Rows nr-w..nr-1 = weight * row nr-w (the row as it was), then row 0 = 0.
__global__ void slab_kill_r_kernel(double* __restrict__ w, const double* __restrict__ wr, int nr,
                                   int nz, int width) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nz) return;
  const double src = w[(int64_t)(nr - width) * nz + j];
  for (int k = 0; k < width; ++k) w[(int64_t)(nr - width + k) * nz + j] = wr[k] * src;
  w[j] = 0.0;
}

// dst (nr, nzi) <- src[:, g:g+nzi] (gather) or dst[:, g:g+nzi] <- src (scatter).
__global__ void slab_columns_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                    int nz, int g, int nzi, int scatter) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nzi) return;
  const int64_t i = blockIdx.y;
  if (scatter)
    dst[i * nz + g + j] = src[i * nzi + j];
  else
    dst[i * nzi + j] = src[i * nz + g + j];
}

// Ghost exchange of up to three fields (blockIdx.z): field[:, :g] =
// field[:, -2g:-g], field[:, -g:] = field[:, g:2g] (disjoint for nz >= 4g).
// Fields with their bit set in `shift` take the reference map's form
// (periodic_boundary_ghost_comm.py:18-32): (x - z_max) + c on the left,
// (x + z_max) - c on the right, c = 2 g dx.
struct SlabFields {
  double* f[3];
};
__global__ void slab_ghost_kernel(SlabFields fs, int nz, int g, int shift, double z_max,
                                  double c) {
  const int q = threadIdx.x;
  if (q >= 2 * g) return;
  double* row = fs.f[blockIdx.z] + (int64_t)blockIdx.x * nz;
  const bool sh = (shift >> blockIdx.z) & 1;
  if (q < g) {
    const double v = row[nz - 2 * g + q];
    row[q] = sh ? (v - z_max) + c : v;
  } else {
    const double v = row[q];
    row[nz - g + (q - g)] = sh ? (v + z_max) - c : v;
  }
}

__global__ void slab_velocity_max_kernel(const double* __restrict__ uz,
                                         const double* __restrict__ ur, int64_t n,
                                         double* __restrict__ slots) {
  __shared__ double wm[4];
  double m = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    m = fmax(m, fabs(uz[k]) + fabs(ur[k]));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
    atomicMax(reinterpret_cast<unsigned long long*>(slots + (blockIdx.x % kMaxSlots) * kSlotStride),
              static_cast<unsigned long long>(__double_as_longlong(m)));
  }
}

// One block of 64: dt = min(0.9 dx^2 / 4 / nu, CFL dx / (max + eps)), the
// derived scalars, U_0 cos(omega t), cleared slots.
__global__ void slab_dt_kernel(double* __restrict__ sc, double* __restrict__ slots,
                               SlabConfig cfg) {
  double m = __longlong_as_double(static_cast<long long>(
      __hip_atomic_load(reinterpret_cast<unsigned long long*>(slots + threadIdx.x * kSlotStride),
                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  slots[threadIdx.x * kSlotStride] = 0.0;
  if (threadIdx.x != 0) return;
  const double visc = 0.9 * (cfg.dx * cfg.dx) / 4.0 / cfg.nu;
  const double adv = cfg.CFL * cfg.dx / (m + cfg.eps);
  const double dt = (adv < visc) ? adv : visc;
  sc[kSlMaxU] = m;
  sc[kSlDt] = dt;
  sc[kSlDtByDx] = dt / cfg.dx;
  sc[kSlHalfNuDt] = 0.5 * cfg.nu * dt;
  sc[kSlNuDt] = cfg.nu * dt;
  sc[kSlU] = cfg.U0 * cos(cfg.omega * sc[kSlT]);
}

// brinkmann_penalize(lam, dt, chi, U, 0, u_upen, u_r_upen, u_z, u_r) with
// u_upen = u: u <- (u + lam dt chi U) / (1 + lam dt chi); du = u_new - u_old.
__global__ void slab_penalize_kernel(double* __restrict__ uz, double* __restrict__ ur,
                                     double* __restrict__ duz, double* __restrict__ dur,
                                     const double* __restrict__ chi,
                                     const double* __restrict__ sc, double lam, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double ldc = (lam * sc[kSlDt]) * chi[k];
  const double den = 1.0 + ldc;
  const double gz = uz[k], gr = ur[k];
  const double pz = (gz + ldc * sc[kSlU]) / den;
  const double pr = (gr + ldc * 0.0) / den;
  uz[k] = pz;
  ur[k] = pr;
  duz[k] = pz - gz;
  dur[k] = pr - gr;
}

// omega += p with p the interior curl of (du_z, du_r) and 0 on the border
// (the reference adds a whole penal-vorticity field whose border stays 0).
__global__ void slab_add_curl_kernel(double* __restrict__ w, const double* __restrict__ duz,
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

__global__ void slab_finalize_kernel(double* __restrict__ sc) {
  if (threadIdx.x != 0) return;
  sc[kSlT] = sc[kSlT] + sc[kSlDt];
  sc[kSlIt] = sc[kSlIt] + 1.0;
}

__global__ void slab_zero_kernel(double* __restrict__ p, int64_t n, int64_t stride) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k * stride] = 0.0;
}

// ------------------------------------------------------------------ stepper

struct SlabStepper {
  SlabConfig cfg{};
  SlabBuffers buf{};
  StokesPlan* plan = nullptr;
  std::vector<void*> allocs;
  double *rhs = nullptr, *psi_in = nullptr, *duz = nullptr, *dur = nullptr;
  double *w1 = nullptr, *tmp = nullptr, *slots = nullptr;
  hipGraph_t graph = nullptr, chunk_graph = nullptr;
  hipGraphExec_t graph_exec = nullptr, chunk_exec = nullptr;
  hipStream_t graph_stream = nullptr;
  static constexpr int kChunkSteps = 16;

  hipError_t alloc(double** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, sizeof(double) * n);
    if (e != hipSuccess) return e;
    allocs.push_back(q);
    *p = static_cast<double*>(q);
    return hipSuccess;
  }

  hipError_t init(const SlabConfig& c, const SlabBuffers& b) {
    cfg = c;
    buf = b;
    plan = new (std::nothrow) StokesPlan();
    if (!plan) return hipErrorOutOfMemory;
    const int nzi = cfg.nz - 2 * cfg.ghost;
    hipError_t e = plan->init(cfg.nr, nzi, cfg.dx, AXS_BC_NEUMANN_R_PERIODIC_Z);
    if (e != hipSuccess) return e;
    const size_t n = (size_t)cfg.nr * cfg.nz, ni = (size_t)cfg.nr * nzi;
    if ((e = alloc(&rhs, ni)) != hipSuccess) return e;
    if ((e = alloc(&psi_in, ni)) != hipSuccess) return e;
    for (double** f : {&duz, &dur, &w1, &tmp})
      if ((e = alloc(f, n)) != hipSuccess) return e;
    if ((e = alloc(&slots, kSlotsPerCase)) != hipSuccess) return e;
    for (double* f : {duz, dur, w1, tmp}) {
      slab_zero_kernel<<<ceil_div((int64_t)n, 256), 256>>>(f, (int64_t)n, 1);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    slab_zero_kernel<<<1, 64>>>(slots, kMaxSlots, kSlotStride);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return hipDeviceSynchronize();
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

  ~SlabStepper() {
    drop_graphs();
    for (void* p : allocs) (void)hipFree(p);
    delete plan;
  }

  hipError_t ghost(double* a, double* b, double* c, int nf, hipStream_t s) {
    slab_ghost_kernel<<<dim3(cfg.nr, 1, nf), 64, 0, s>>>(SlabFields{{a, b, c}}, cfg.nz, cfg.ghost,
                                                         0, 0.0, 0.0);
    return hipGetLastError();
  }

  hipError_t enqueue_step(hipStream_t s) {
    const int nr = cfg.nr, nz = cfg.nz, g = cfg.ghost, nzi = nz - 2 * g;
    const int64_t n = (int64_t)nr * nz;
    const dim3 g2(ceil_div(nz, 256), nr), gi(ceil_div(nzi, 256), nr);
    double* sc = buf.scalars;
    hipError_t e;
#define AXS_TRY_STEP(...)                            \
  do {                                               \
    if ((e = (__VA_ARGS__)) != hipSuccess) return e; \
  } while (0)
#define AXS_TRY_LAUNCH()                                 \
  do {                                                   \
    if ((e = hipGetLastError()) != hipSuccess) return e; \
  } while (0)
    slab_kill_r_kernel<<<ceil_div(nz, 256), 256, 0, s>>>(buf.vorticity, buf.kill_wr, nr, nz,
                                                         cfg.kill_width);
    AXS_TRY_LAUNCH();
    slab_columns_kernel<<<gi, 256, 0, s>>>(buf.vorticity, rhs, nz, g, nzi, 0);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(launch_rfft_rows(*plan, rhs, plan->spec, plan->rvec, nr, s, false));
    AXS_TRY_STEP(launch_tridiag<double>(*plan, plan->spec, s));
    AXS_TRY_STEP(launch_rfft_rows(*plan, plan->spec, psi_in, nullptr, nr, s, true));
    slab_columns_kernel<<<gi, 256, 0, s>>>(psi_in, buf.psi, nz, g, nzi, 1);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(ghost(buf.psi, nullptr, nullptr, 1, s));
    AXS_TRY_STEP(launch_velocity_from_psi<double>(buf.u_z, buf.u_r, buf.psi, buf.r, nr, nz,
                                                  cfg.dx, s));
    const int mb = (int)std::min<int64_t>(ceil_div(n, 256), 1024);
    slab_velocity_max_kernel<<<mb, 256, 0, s>>>(buf.u_z, buf.u_r, n, slots);
    AXS_TRY_LAUNCH();
    slab_dt_kernel<<<1, 64, 0, s>>>(sc, slots, cfg);
    AXS_TRY_LAUNCH();
    slab_penalize_kernel<<<ceil_div(n, 256), 256, 0, s>>>(buf.u_z, buf.u_r, duz, dur,
                                                          buf.char_func, sc, cfg.lam, n);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(ghost(dur, duz, nullptr, 2, s));
    slab_add_curl_kernel<<<g2, 256, 0, s>>>(buf.vorticity, duz, dur, nr, nz, cfg.dx);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(ghost(buf.u_z, buf.u_r, buf.vorticity, 3, s));
    AXS_TRY_STEP(launch_eno3_conservative<double>(buf.vorticity, buf.u_z, buf.u_r, w1, nr, nz,
                                                  Scalar<double>{0.0, sc + kSlDtByDx}, -1.0, s));
    AXS_TRY_STEP(ghost(w1, nullptr, nullptr, 1, s));
    AXS_TRY_STEP(launch_rk2_stage1<double, RRow<double>>(
        w1, tmp, RRow<double>{buf.r}, nr, nz, Scalar<double>{0.0, sc + kSlHalfNuDt}, cfg.dx, s));
    AXS_TRY_STEP(ghost(tmp, nullptr, nullptr, 1, s));
    AXS_TRY_STEP(launch_rk2_stage2<double, RRow<double>>(w1, tmp, buf.vorticity,
                                                         RRow<double>{buf.r}, nr, nz,
                                                         Scalar<double>{0.0, sc + kSlNuDt},
                                                         cfg.dx, s));
    slab_finalize_kernel<<<1, 64, 0, s>>>(sc);
    AXS_TRY_LAUNCH();
#undef AXS_TRY_STEP
#undef AXS_TRY_LAUNCH
    return hipSuccess;
  }

  hipError_t capture(hipStream_t s) {
    drop_graphs();
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

  hipError_t run(int nsteps, bool use_graph, hipStream_t s) {
    hipError_t e;
    if (!use_graph) {
      for (int q = 0; q < nsteps; ++q)
        if ((e = enqueue_step(s)) != hipSuccess) return e;
      return hipSuccess;
    }
    if (graph_stream != s || !graph_exec)
      if ((e = capture(s)) != hipSuccess) return e;
    int q = 0;
    for (; q + kChunkSteps <= nsteps; q += kChunkSteps)
      if ((e = hipGraphLaunch(chunk_exec, s)) != hipSuccess) return e;
    for (; q < nsteps; ++q)
      if ((e = hipGraphLaunch(graph_exec, s)) != hipSuccess) return e;
    return hipSuccess;
  }
};


// ================================================================ soft slab
//
// Device-resident timestep of the periodic soft-slab driver
// (examples/PeriodicSoftSlab/periodic_soft_slab.py:214-396) on the same
// grid layout: an elastic layer around the axis (its reference map eta1,
// eta2 advected, masked to the solid and extrapolated by the static PDE
// scheme of bounded_static_PDE_extrapolation.py, its blended stress added to
// omega) driven by the oscillating annular wall through the fluid.  Order:
//   velocity max, dt    dt = min(CFL dx / sqrt(G / rho), 0.9 dx^2 / 4 / nu,
//                       CFL dx / (max + eps)) from the previous step's
//                       velocity, clipped to the sample timer and t_end;
//                       the timer resets when it has run out (:216-238)
//   kill_r, psi         as the rigid slab
//   ghost(psi), velocity, sample (t and u_z of column nz / 2)
//   ghost(u, eta), refmap ENO3 -> e1n, e2n   (advect_refmap_via_eno3_periodic)
//   penalise, ghost(du), curl                 (the wall, as the rigid slab)
//   mask                eta = (solid_char_func > 0.5) * e_n (:366-367)
//   extrapolate         both components, one workgroup each (below)
//   ghost(eta), sigma (blended), ghost(tau), tau_z / tau_r, ghost, curl
//                       (solid_sigma_periodic, :372-386, div_tau periodic)
//   ghost(u, omega), ENO3, ghost, RK2 stage 1, ghost, RK2 stage 2
//   finalize            t += dt, timer += dt, it += 1

enum : int {
  kSfT = 0,     // time
  kSfDt,        // dt of the current step
  kSfDtByDx,    // dt / dx
  kSfHalfNuDt,  // 0.5 nu dt
  kSfNuDt,      // nu dt
  kSfU,         // wall velocity V_wall cos(omega t)
  kSfMaxU,      // max(|u_z| + |u_r|) the step's dt used
  kSfIt,        // step counter
  kSfTimer,     // sample timer
  kSfSamples,   // samples taken
  kSfJacobi,    // Jacobi iterations of the last step's extrapolation (4 loops)
  kSfCapped,    // loops stopped at the iteration bound (all steps)
  kSfSampleNow, // 1 when the current step samples
  kSfCount
};
static_assert(kSfCount <= AXS_RIGID_SCALARS, "soft-slab scalar slots exceed the array");

struct SoftSlabConfig {
  int nr, nz, ghost, kill_width, sample_rows, sample_cap, max_jacobi;
  double dx, nu, G, rho, lam, V_wall, omega, CFL, eps, z_max, tol, band, freq_limit, t_end;
};

struct SoftSlabBuffers {
  double *vorticity, *psi, *u_z, *u_r, *eta1, *eta2, *solid_phi, *scalars, *samples;
  const double *solid_chi, *wall_chi, *r, *kill_wr;
};

// The extrapolation's fixed data: solid_phi does not change in this driver,
// so the box (rows r0 .. r0 + bR, columns z0 .. z0 + bZ), the box rows and
// columns the result is restored to, and the per-cell zones and normals
// (:152-183) are set up once on the host.  cf holds ten arrays over the box
// interior ((bR - 4) x (bZ - 4)), in this order.
enum : int { kCfInside = 0, kCfZone, kCfNrp, kCfNrm, kCfNzp, kCfNzm, kCfDenom, kCfNr, kCfNz,
             kCfRhs0, kCfCount };
struct SlabExtrapPlan {
  int nr, nz, g;
  int r0, z0, bR, bZ;
  int rs, re, zs, ze;  // restored box rows [rs, re), columns [zs, ze)
  int max_jacobi;
  double two_dx, z_max, c, tol;
  const double* cf;
  double* wk;  // per field: four box arrays and one interior array
};

__global__ void softslab_dt_kernel(double* __restrict__ sc, double* __restrict__ slots,
                                   SoftSlabConfig cfg) {
  double m = __longlong_as_double(static_cast<long long>(
      __hip_atomic_load(reinterpret_cast<unsigned long long*>(slots + threadIdx.x * kSlotStride),
                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  slots[threadIdx.x * kSlotStride] = 0.0;
  if (threadIdx.x != 0) return;
  // Python's min over the three bounds in their order (:216-220)
  double dt = cfg.CFL * cfg.dx / sqrt(cfg.G / cfg.rho);
  const double visc = 0.9 * (cfg.dx * cfg.dx) / 4.0 / cfg.nu;
  const double adv = cfg.CFL * cfg.dx / (m + cfg.eps);
  if (visc < dt) dt = visc;
  if (adv < dt) dt = adv;
  const double timer = sc[kSfTimer], t = sc[kSfT];
  if (timer + dt > cfg.freq_limit) dt = cfg.freq_limit - timer;
  if (t + dt > cfg.t_end) dt = cfg.t_end - t;
  const bool sample = timer >= cfg.freq_limit;
  if (sample) sc[kSfTimer] = 0.0;
  sc[kSfSampleNow] = sample ? 1.0 : 0.0;
  sc[kSfMaxU] = m;
  sc[kSfDt] = dt;
  sc[kSfDtByDx] = dt / cfg.dx;
  sc[kSfHalfNuDt] = 0.5 * cfg.nu * dt;
  sc[kSfNuDt] = cfg.nu * dt;
  sc[kSfU] = cfg.V_wall * cos(cfg.omega * t);
  sc[kSfJacobi] = 0.0;
}

// The driver's sample (:229-238): t, then u_z of rows < rows in column nz / 2.
__global__ void softslab_sample_kernel(const double* __restrict__ uz, double* __restrict__ sc,
                                       double* __restrict__ samples, int nz, int rows, int cap) {
  if (sc[kSfSampleNow] == 0.0) return;
  const int k = (int)sc[kSfSamples];
  if (k < cap) {
    double* out = samples + (int64_t)k * (rows + 1);
    for (int i = threadIdx.x; i < rows; i += blockDim.x) out[1 + i] = uz[(int64_t)i * nz + nz / 2];
    if (threadIdx.x == 0) out[0] = sc[kSfT];
  }
  __syncthreads();
  if (threadIdx.x == 0) sc[kSfSamples] = sc[kSfSamples] + 1.0;
}

// eta = inside_solid * e_n with inside_solid = solid_char_func > 0.5.
__global__ void softslab_mask_kernel(double* __restrict__ e1, double* __restrict__ e2,
                                     const double* __restrict__ n1, const double* __restrict__ n2,
                                     const double* __restrict__ chi, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double in = chi[k] > 0.5 ? 1.0 : 0.0;
  e1[k] = in * n1[k];
  e2[k] = in * n2[k];
}

__device__ __forceinline__ double block_sum_1024(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();  // red may still be read from the call before
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int q = 0; q < nw; ++q) t += red[q];
  return t;
}

// Jacobi iterations of n . grad(S) = rhs on the zone (:185-218), S fixed
// inside the solid: box arrays S and T (same margins), the interior's
// right-hand side rhs (box-indexed when rhs_box), `orig` interior scratch.
// Runs while the L2 norm of an iteration's change is above tol, at most
// max_jacobi times; returns the array holding the result.
__device__ double* softslab_jacobi(double* S, double* T, const double* rhs, bool rhs_box,
                                   double* orig, const SlabExtrapPlan& P, int* its_out,
                                   double* red) {
  const int tid = threadIdx.x, nth = blockDim.x;
  const int bZ = P.bZ, iZ = bZ - 4, ni = (P.bR - 4) * iZ;
  const double* cf = P.cf;
  for (int k = tid; k < ni; k += nth) {
    const int e = (k / iZ + 2) * bZ + k % iZ + 2;
    orig[k] = S[e] * cf[kCfInside * ni + k];
  }
  int its = 0;
  for (;;) {
    double acc = 0.0;
    for (int k = tid; k < ni; k += nth) {
      const int e = (k / iZ + 2) * bZ + k % iZ + 2;
      const double num_r = cf[kCfNrp * ni + k] * (4.0 * S[e - bZ] - S[e - 2 * bZ]) -
                           cf[kCfNrm * ni + k] * (4.0 * S[e + bZ] - S[e + 2 * bZ]);
      const double num_z = cf[kCfNzp * ni + k] * (4.0 * S[e - 1] - S[e - 2]) -
                           cf[kCfNzm * ni + k] * (4.0 * S[e + 1] - S[e + 2]);
      const double rv = rhs_box ? rhs[e] : rhs[k];
      const double v = orig[k] + cf[kCfZone * ni + k] * (P.two_dx * rv + num_r + num_z) /
                                     cf[kCfDenom * ni + k];
      T[e] = v;
      const double d = v - S[e];
      acc += d * d;
    }
    const double res = sqrt(block_sum_1024(acc, red));  // every thread: the same value
    ++its;
    double* t = S;
    S = T;
    T = t;
    if (!(res > P.tol) || its >= P.max_jacobi) break;
  }
  *its_out = its;
  return S;
}

// bounded_static_PDE_extrapolation.py:58-123 for eta1 (block 0) or eta2
// (block 1): ghost exchange (the driver passes the reference-map form for
// both), the box, grad_eta_n inside the solid, the two Jacobi loops, then
// eta = 0 with the restored box (:220-235).
__global__ void __launch_bounds__(1024) softslab_extrap_kernel(double* __restrict__ eta1,
                                                               double* __restrict__ eta2,
                                                               SlabExtrapPlan P,
                                                               double* __restrict__ sc) {
  __shared__ double red[16];
  double* eta = blockIdx.x ? eta2 : eta1;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int nr = P.nr, nz = P.nz, g = P.g;
  const int bR = P.bR, bZ = P.bZ, nb = bR * bZ, iZ = bZ - 4, ni = (bR - 4) * iZ;
  double* W = P.wk + (size_t)blockIdx.x * (4 * (size_t)nb + ni);
  double *S0 = W, *S1 = W + nb, *B0 = W + 2 * nb, *B1 = W + 3 * nb, *orig = W + 4 * nb;
  for (int e = tid; e < nr * 2 * g; e += nth) {
    double* row = eta + (int64_t)(e / (2 * g)) * nz;
    const int q = e % (2 * g);
    if (q < g)
      row[q] = (row[nz - 2 * g + q] - P.z_max) + P.c;
    else
      row[nz - g + (q - g)] = (row[q] + P.z_max) - P.c;
  }
  __syncthreads();
  for (int e = tid; e < nb; e += nth) {
    const double v = eta[(int64_t)(P.r0 + e / bZ) * nz + P.z0 + e % bZ];
    B0[e] = v;
    B1[e] = v;
    S0[e] = v * 0.0;  // grad_eta_n = bounded_eta * 0
    S1[e] = v * 0.0;
  }
  __syncthreads();
  for (int k = tid; k < ni; k += nth) {
    const int e = (k / iZ + 2) * bZ + k % iZ + 2;
    const double gr = (B0[e + bZ] - B0[e - bZ]) / P.two_dx;
    const double gz = (B0[e + 1] - B0[e - 1]) / P.two_dx;
    const double v =
        P.cf[kCfInside * ni + k] * (P.cf[kCfNr * ni + k] * gr + P.cf[kCfNz * ni + k] * gz);
    S0[e] = v;
    S1[e] = v;
  }
  __syncthreads();
  int its1 = 0, its2 = 0;
  const double* gn = softslab_jacobi(S0, S1, P.cf + kCfRhs0 * ni, false, orig, P, &its1, red);
  __syncthreads();
  const double* res = softslab_jacobi(B0, B1, gn, true, orig, P, &its2, red);
  __syncthreads();
  for (int c = tid; c < nr * nz; c += nth) {
    const int a = c / nz - P.r0, b = c % nz - P.z0;
    eta[c] = (a >= P.rs && a < P.re && b >= P.zs && b < P.ze) ? res[a * bZ + b] : eta[c] * 0.0;
  }
  if (tid == 0) {
    atomicAdd(sc + kSfJacobi, (double)(its1 + its2));
    const int capped = (its1 >= P.max_jacobi) + (its2 >= P.max_jacobi);
    if (capped) atomicAdd(sc + kSfCapped, (double)capped);
  }
}

__global__ void softslab_finalize_kernel(double* __restrict__ sc) {
  if (threadIdx.x != 0) return;
  sc[kSfT] = sc[kSfT] + sc[kSfDt];
  sc[kSfTimer] = sc[kSfTimer] + sc[kSfDt];
  sc[kSfIt] = sc[kSfIt] + 1.0;
}

// numpy.maximum(x, 0) / numpy.minimum(x, 0): x when the comparison holds or
// x is NaN, else 0.
static double np_max0(double x) { return (x >= 0.0 || std::isnan(x)) ? x : 0.0; }
static double np_min0(double x) { return (x <= 0.0 || std::isnan(x)) ? x : 0.0; }

struct SoftSlabStepper {
  SoftSlabConfig cfg{};
  SoftSlabBuffers buf{};
  StokesPlan* plan = nullptr;
  SlabExtrapPlan xp{};
  std::vector<void*> allocs;
  double *rhs = nullptr, *psi_in = nullptr, *duz = nullptr, *dur = nullptr;
  double *w1 = nullptr, *tmp = nullptr, *slots = nullptr, *e1n = nullptr, *e2n = nullptr;
  double *s11 = nullptr, *s12 = nullptr, *s22 = nullptr, *g1z = nullptr, *g1r = nullptr;
  double *g2z = nullptr, *g2r = nullptr, *tz = nullptr, *tr = nullptr;
  double *cf = nullptr, *wk = nullptr;
  hipGraph_t graph = nullptr, chunk_graph = nullptr;
  hipGraphExec_t graph_exec = nullptr, chunk_exec = nullptr;
  hipStream_t graph_stream = nullptr;
  static constexpr int kChunkSteps = 16;

  hipError_t alloc(double** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, sizeof(double) * n);
    if (e != hipSuccess) return e;
    allocs.push_back(q);
    *p = static_cast<double*>(q);
    return hipSuccess;
  }

  // :125-183 on solid_phi (host copy, ghost columns exchanged); false when
  // the box is too small (:149-150)
  bool plan_extrapolation(const std::vector<double>& phi, std::vector<double>& h_cf) {
    const int nr = cfg.nr, nz = cfg.nz;
    int r_lo = -1, r_hi = -1, z_lo = -1, z_hi = -1;
    for (int i = 0; i < nr; ++i)
      for (int j = 0; j < nz; ++j)
        if (phi[(size_t)i * nz + j] + cfg.band >= 0) {
          if (r_lo < 0) r_lo = i;
          r_hi = i;
          if (z_lo < 0 || j < z_lo) z_lo = j;
          if (j > z_hi) z_hi = j;
        }
    if (r_lo < 0) return false;
    int r0 = r_lo, r1 = r_hi + 1, z0 = z_lo, z1 = z_hi + 1;
    bool rsb = true, reb = true, zsb = true, zeb = true;
    if (r0 >= 2) { r0 -= 2; rsb = false; }
    if (r1 <= nr - 2) { r1 += 2; reb = false; }
    if (z0 >= 2) { z0 -= 2; zsb = false; }
    if (z1 <= nz - 2) { z1 += 2; zeb = false; }
    if (std::min(r1 - r0, z1 - z0) < 6) return false;
    const int bR = r1 - r0, bZ = z1 - z0, iR = bR - 4, iZ = bZ - 4, ni = iR * iZ;
    xp.nr = nr;
    xp.nz = nz;
    xp.g = cfg.ghost;
    xp.r0 = r0;
    xp.z0 = z0;
    xp.bR = bR;
    xp.bZ = bZ;
    xp.rs = rsb ? 0 : 2;
    xp.re = reb ? bR : bR - 2;
    xp.zs = zsb ? 0 : 2;
    xp.ze = zeb ? bZ : bZ - 2;
    xp.max_jacobi = cfg.max_jacobi;
    xp.two_dx = 2 * cfg.dx;
    xp.z_max = cfg.z_max;
    xp.c = (double)(2 * cfg.ghost) * cfg.dx;
    xp.tol = cfg.tol;
    auto bphi = [&](int a, int b) { return -phi[(size_t)(r0 + a) * nz + z0 + b]; };
    const double offset = std::sqrt(2.0) * cfg.dx;
    const double eps = std::numeric_limits<double>::epsilon();
    h_cf.assign((size_t)kCfCount * ni, 0.0);
    for (int k = 0; k < ni; ++k) {
      const int a = k / iZ + 2, b = k % iZ + 2;
      const double c = bphi(a, b);
      double n_r = (np_max0(3.0 * c - 4.0 * bphi(a - 1, b) + bphi(a - 2, b)) +
                    np_min0(-3.0 * c + 4.0 * bphi(a + 1, b) - bphi(a + 2, b))) /
                   (2 * cfg.dx);
      double n_z = (np_max0(3.0 * c - 4.0 * bphi(a, b - 1) + bphi(a, b - 2)) +
                    np_min0(-3.0 * c + 4.0 * bphi(a, b + 1) - bphi(a, b + 2))) /
                   (2 * cfg.dx);
      const double mag = std::sqrt(n_r * n_r + n_z * n_z);
      n_r /= mag + eps;
      n_z /= mag + eps;
      const double nrp = (n_r >= 0 ? 1.0 : 0.0) * n_r, nzp = (n_z >= 0 ? 1.0 : 0.0) * n_z;
      h_cf[kCfInside * ni + k] = c <= -offset ? 1.0 : 0.0;
      h_cf[kCfZone * ni + k] = (c > -offset && c < cfg.band) ? 1.0 : 0.0;
      h_cf[kCfNrp * ni + k] = nrp;
      h_cf[kCfNrm * ni + k] = n_r - nrp;
      h_cf[kCfNzp * ni + k] = nzp;
      h_cf[kCfNzm * ni + k] = n_z - nzp;
      h_cf[kCfDenom * ni + k] = 3.0 * (std::fabs(n_r) + std::fabs(n_z)) + eps;
      h_cf[kCfNr * ni + k] = n_r;
      h_cf[kCfNz * ni + k] = n_z;
      h_cf[kCfRhs0 * ni + k] = c * 0.0;
    }
    return true;
  }

  hipError_t init(const SoftSlabConfig& c, const SoftSlabBuffers& b, bool* box_ok) {
    cfg = c;
    buf = b;
    *box_ok = true;
    plan = new (std::nothrow) StokesPlan();
    if (!plan) return hipErrorOutOfMemory;
    const int nr = cfg.nr, nz = cfg.nz, g = cfg.ghost, nzi = nz - 2 * g;
    hipError_t e = plan->init(nr, nzi, cfg.dx, AXS_BC_NEUMANN_R_PERIODIC_Z);
    if (e != hipSuccess) return e;
    const size_t n = (size_t)nr * nz, ni = (size_t)nr * nzi;
    if ((e = alloc(&rhs, ni)) != hipSuccess) return e;
    if ((e = alloc(&psi_in, ni)) != hipSuccess) return e;
    for (double** f : {&duz, &dur, &w1, &tmp, &e1n, &e2n, &s11, &s12, &s22, &g1z, &g1r, &g2z,
                       &g2r, &tz, &tr})
      if ((e = alloc(f, n)) != hipSuccess) return e;
    if ((e = alloc(&slots, kSlotsPerCase)) != hipSuccess) return e;
    // scratch the reference allocates zeroed and never fully rewrites
    for (double* f : {duz, dur, w1, tmp, e1n, e2n, s11, s12, s22, g1z, g1r, g2z, g2r, tz, tr}) {
      slab_zero_kernel<<<ceil_div((int64_t)n, 256), 256>>>(f, (int64_t)n, 1);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    slab_zero_kernel<<<1, 64>>>(slots, kMaxSlots, kSlotStride);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // solid_phi's ghost columns (the extrapolation exchanges them first)
    slab_ghost_kernel<<<dim3(nr, 1, 1), 64>>>(SlabFields{{buf.solid_phi, nullptr, nullptr}}, nz, g,
                                              0, 0.0, 0.0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipDeviceSynchronize()) != hipSuccess) return e;
    std::vector<double> phi(n), h_cf;
    if ((e = hipMemcpy(phi.data(), buf.solid_phi, sizeof(double) * n, hipMemcpyDeviceToHost)) !=
        hipSuccess)
      return e;
    if (!plan_extrapolation(phi, h_cf)) {
      *box_ok = false;
      return hipSuccess;
    }
    if ((e = alloc(&cf, h_cf.size())) != hipSuccess) return e;
    if ((e = hipMemcpy(cf, h_cf.data(), sizeof(double) * h_cf.size(), hipMemcpyHostToDevice)) !=
        hipSuccess)
      return e;
    const size_t nb = (size_t)xp.bR * xp.bZ, nbi = (size_t)(xp.bR - 4) * (xp.bZ - 4);
    if ((e = alloc(&wk, 2 * (4 * nb + nbi))) != hipSuccess) return e;
    xp.cf = cf;
    xp.wk = wk;
    return hipDeviceSynchronize();
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

  ~SoftSlabStepper() {
    drop_graphs();
    for (void* p : allocs) (void)hipFree(p);
    delete plan;
  }

  hipError_t ghost(double* a, double* b, double* c, int nf, int shift, hipStream_t s) {
    slab_ghost_kernel<<<dim3(cfg.nr, 1, nf), 64, 0, s>>>(
        SlabFields{{a, b, c}}, cfg.nz, cfg.ghost, shift, cfg.z_max,
        (double)(2 * cfg.ghost) * cfg.dx);
    return hipGetLastError();
  }

  hipError_t enqueue_step(hipStream_t s) {
    const int nr = cfg.nr, nz = cfg.nz, g = cfg.ghost, nzi = nz - 2 * g;
    const int64_t n = (int64_t)nr * nz;
    const dim3 g2(ceil_div(nz, 256), nr), gi(ceil_div(nzi, 256), nr);
    double* sc = buf.scalars;
    hipError_t e;
#define AXS_TRY_STEP(...)                            \
  do {                                               \
    if ((e = (__VA_ARGS__)) != hipSuccess) return e; \
  } while (0)
#define AXS_TRY_LAUNCH()                                 \
  do {                                                   \
    if ((e = hipGetLastError()) != hipSuccess) return e; \
  } while (0)
    const int mb = (int)std::min<int64_t>(ceil_div(n, 256), 1024);
    slab_velocity_max_kernel<<<mb, 256, 0, s>>>(buf.u_z, buf.u_r, n, slots);
    AXS_TRY_LAUNCH();
    softslab_dt_kernel<<<1, 64, 0, s>>>(sc, slots, cfg);
    AXS_TRY_LAUNCH();
    slab_kill_r_kernel<<<ceil_div(nz, 256), 256, 0, s>>>(buf.vorticity, buf.kill_wr, nr, nz,
                                                         cfg.kill_width);
    AXS_TRY_LAUNCH();
    slab_columns_kernel<<<gi, 256, 0, s>>>(buf.vorticity, rhs, nz, g, nzi, 0);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(launch_rfft_rows(*plan, rhs, plan->spec, plan->rvec, nr, s, false));
    AXS_TRY_STEP(launch_tridiag<double>(*plan, plan->spec, s));
    AXS_TRY_STEP(launch_rfft_rows(*plan, plan->spec, psi_in, nullptr, nr, s, true));
    slab_columns_kernel<<<gi, 256, 0, s>>>(psi_in, buf.psi, nz, g, nzi, 1);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(ghost(buf.psi, nullptr, nullptr, 1, 0, s));
    AXS_TRY_STEP(launch_velocity_from_psi<double>(buf.u_z, buf.u_r, buf.psi, buf.r, nr, nz,
                                                  cfg.dx, s));
    softslab_sample_kernel<<<1, 256, 0, s>>>(buf.u_z, sc, buf.samples, nz, cfg.sample_rows,
                                             cfg.sample_cap);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(ghost(buf.u_z, buf.u_r, nullptr, 2, 0, s));
    AXS_TRY_STEP(ghost(buf.eta1, buf.eta2, nullptr, 2, 1, s));
    AXS_TRY_STEP(launch_eno3_refmap<double>(buf.eta1, buf.eta2, buf.u_z, buf.u_r, e1n, e2n, nr,
                                            nz, Scalar<double>{0.0, sc + kSfDtByDx}, s));
    slab_penalize_kernel<<<ceil_div(n, 256), 256, 0, s>>>(buf.u_z, buf.u_r, duz, dur,
                                                          buf.wall_chi, sc, cfg.lam, n);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(ghost(dur, duz, nullptr, 2, 0, s));
    slab_add_curl_kernel<<<g2, 256, 0, s>>>(buf.vorticity, duz, dur, nr, nz, cfg.dx);
    AXS_TRY_LAUNCH();
    softslab_mask_kernel<<<ceil_div(n, 256), 256, 0, s>>>(buf.eta1, buf.eta2, e1n, e2n,
                                                          buf.solid_chi, n);
    AXS_TRY_LAUNCH();
    softslab_extrap_kernel<<<2, 1024, 0, s>>>(buf.eta1, buf.eta2, xp, sc);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(ghost(buf.eta1, buf.eta2, nullptr, 2, 1, s));
    AXS_TRY_STEP(launch_solid_sigma<double>(s11, s12, s22, cfg.G, cfg.dx, buf.eta1, buf.eta2, g1z,
                                            g1r, g2z, g2r, nr, nz, s, buf.solid_chi));
    AXS_TRY_STEP(ghost(s11, s12, s22, 3, 0, s));
    AXS_TRY_STEP(launch_div_tau_stage1<double>(tz, tr, s11, s12, s22, buf.r, nr, nz, cfg.dx, s));
    AXS_TRY_STEP(ghost(tr, tz, nullptr, 2, 0, s));
    AXS_TRY_STEP(launch_div_tau_stage2<double>(buf.vorticity, tz, tr, nr, nz,
                                               Scalar<double>{0.0, sc + kSfDt}, cfg.dx, s));
    AXS_TRY_STEP(ghost(buf.u_z, buf.u_r, buf.vorticity, 3, 0, s));
    AXS_TRY_STEP(launch_eno3_conservative<double>(buf.vorticity, buf.u_z, buf.u_r, w1, nr, nz,
                                                  Scalar<double>{0.0, sc + kSfDtByDx}, -1.0, s));
    AXS_TRY_STEP(ghost(w1, nullptr, nullptr, 1, 0, s));
    AXS_TRY_STEP(launch_rk2_stage1<double, RRow<double>>(
        w1, tmp, RRow<double>{buf.r}, nr, nz, Scalar<double>{0.0, sc + kSfHalfNuDt}, cfg.dx, s));
    AXS_TRY_STEP(ghost(tmp, nullptr, nullptr, 1, 0, s));
    AXS_TRY_STEP(launch_rk2_stage2<double, RRow<double>>(w1, tmp, buf.vorticity,
                                                         RRow<double>{buf.r}, nr, nz,
                                                         Scalar<double>{0.0, sc + kSfNuDt},
                                                         cfg.dx, s));
    softslab_finalize_kernel<<<1, 64, 0, s>>>(sc);
    AXS_TRY_LAUNCH();
#undef AXS_TRY_STEP
#undef AXS_TRY_LAUNCH
    return hipSuccess;
  }

  hipError_t capture(hipStream_t s) {
    drop_graphs();
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

  hipError_t run(int nsteps, bool use_graph, hipStream_t s) {
    hipError_t e;
    if (!use_graph) {
      for (int q = 0; q < nsteps; ++q)
        if ((e = enqueue_step(s)) != hipSuccess) return e;
      return hipSuccess;
    }
    if (graph_stream != s || !graph_exec)
      if ((e = capture(s)) != hipSuccess) return e;
    int q = 0;
    for (; q + kChunkSteps <= nsteps; q += kChunkSteps)
      if ((e = hipGraphLaunch(chunk_exec, s)) != hipSuccess) return e;
    for (; q < nsteps; ++q)
      if ((e = hipGraphLaunch(graph_exec, s)) != hipSuccess) return e;
    return hipSuccess;
  }
};

}  // namespace axs

using namespace axs;

struct axs_slab_stepper {
  SlabStepper impl;
};

struct axs_soft_slab_stepper {
  SoftSlabStepper impl;
};

extern "C" {

axs_status axs_slab_stepper_create(const axs_slab_config* c, const axs_slab_buffers* b,
                                   axs_slab_stepper** out) {
  AXS_CHECK_ARG(c && b && out, "null pointer");
  *out = nullptr;
  AXS_CHECK_ARG(c->ghost >= 1 && c->ghost <= 32 && c->nz - 2 * c->ghost >= 2 &&
                    c->nz >= 4 * c->ghost,
                "need 1 <= ghost <= 32, nz >= 4 ghost and at least 2 interior columns");
  AXS_CHECK_ARG(c->nr >= 4 && c->nr <= 4096, "need 4 <= nr <= 4096");
  AXS_CHECK_ARG(stokes_nz_supported(c->nz - 2 * c->ghost), "interior width not supported");
  AXS_CHECK_ARG(c->kill_width >= 2 && c->kill_width < c->nr, "bad kill_width for the grid");
  AXS_CHECK_ARG(c->dx > 0 && c->nu > 0 && c->CFL > 0 && c->lam >= 0,
                "dx, nu and CFL must be positive, lam non-negative");
  AXS_CHECK_ARG(b->vorticity && b->psi && b->u_z && b->u_r && b->scalars && b->char_func &&
                    b->r && b->kill_wr,
                "null buffer");
  axs_slab_stepper* st = new (std::nothrow) axs_slab_stepper();
  if (!st) {
    set_error("out of host memory");
    return AXS_ERR_ALLOC;
  }
  SlabConfig cfg{c->nr, c->nz, c->ghost, c->kill_width, c->dx, c->nu,
                 c->lam, c->U0, c->omega, c->CFL, c->eps};
  SlabBuffers bb{b->vorticity, b->psi, b->u_z, b->u_r, b->scalars, b->char_func, b->r, b->kill_wr};
  hipError_t e = st->impl.init(cfg, bb);
  if (e != hipSuccess) {
    set_error(std::string("slab stepper: ") + hipGetErrorString(e));
    delete st;
    return AXS_ERR_HIP;
  }
  *out = st;
  return AXS_OK;
}

axs_status axs_slab_stepper_destroy(axs_slab_stepper* st) {
  delete st;
  return AXS_OK;
}

axs_status axs_slab_stepper_run(axs_slab_stepper* st, int nsteps, int use_graph, void* stream) {
  AXS_CHECK_ARG(st, "null stepper");
  AXS_CHECK_ARG(nsteps >= 0, "negative step count");
  AXS_CHECK_ARG(!use_graph || stream, "graph capture needs a non-null stream");
  AXS_HIP(st->impl.run(nsteps, use_graph != 0, as_stream(stream)));
  return AXS_OK;
}

axs_status axs_soft_slab_stepper_create(const axs_soft_slab_config* c,
                                        const axs_soft_slab_buffers* b,
                                        axs_soft_slab_stepper** out) {
  AXS_CHECK_ARG(c && b && out, "null pointer");
  *out = nullptr;
  AXS_CHECK_ARG(c->ghost >= 1 && c->ghost <= 32 && c->nz - 2 * c->ghost >= 2 &&
                    c->nz >= 4 * c->ghost,
                "need 1 <= ghost <= 32, nz >= 4 ghost and at least 2 interior columns");
  AXS_CHECK_ARG(c->nr >= 6 && c->nr <= 4096, "need 6 <= nr <= 4096");
  AXS_CHECK_ARG(stokes_nz_supported(c->nz - 2 * c->ghost), "interior width not supported");
  AXS_CHECK_ARG(c->kill_width >= 2 && c->kill_width < c->nr, "bad kill_width for the grid");
  AXS_CHECK_ARG(c->dx > 0 && c->nu > 0 && c->G > 0 && c->rho > 0 && c->CFL > 0 && c->lam >= 0,
                "dx, nu, G, rho and CFL must be positive, lam non-negative");
  AXS_CHECK_ARG(c->extrap_tol > 0 && c->extrap_band > 0 && c->max_extrap_iterations >= 1,
                "extrap_tol, extrap_band and max_extrap_iterations must be positive");
  AXS_CHECK_ARG(c->freq_limit > 0, "freq_limit must be positive");
  AXS_CHECK_ARG(c->sample_rows >= 0 && c->sample_rows <= c->nr && c->sample_cap >= 0,
                "bad sample shape");
  AXS_CHECK_ARG(c->sample_cap == 0 || b->samples, "samples buffer missing");
  AXS_CHECK_ARG(b->vorticity && b->psi && b->u_z && b->u_r && b->eta1 && b->eta2 &&
                    b->solid_phi && b->scalars && b->solid_char_func && b->wall_char_func &&
                    b->r && b->kill_wr,
                "null buffer");
  axs_soft_slab_stepper* st = new (std::nothrow) axs_soft_slab_stepper();
  if (!st) {
    set_error("out of host memory");
    return AXS_ERR_ALLOC;
  }
  SoftSlabConfig cfg{c->nr, c->nz, c->ghost, c->kill_width, c->sample_rows, c->sample_cap,
                     c->max_extrap_iterations, c->dx, c->nu, c->G, c->rho, c->lam, c->V_wall,
                     c->omega, c->CFL, c->eps, c->z_max, c->extrap_tol, c->extrap_band,
                     c->freq_limit, c->t_end};
  SoftSlabBuffers bb{b->vorticity, b->psi, b->u_z, b->u_r, b->eta1, b->eta2, b->solid_phi,
                     b->scalars, b->samples, b->solid_char_func, b->wall_char_func, b->r,
                     b->kill_wr};
  bool box_ok = true;
  hipError_t e = st->impl.init(cfg, bb, &box_ok);
  if (e != hipSuccess) {
    set_error(std::string("soft slab stepper: ") + hipGetErrorString(e));
    delete st;
    return AXS_ERR_HIP;
  }
  if (!box_ok) {
    set_error("soft slab stepper: the extrapolation box needs at least 6 cells each way");
    delete st;
    return AXS_ERR_INVALID;
  }
  *out = st;
  return AXS_OK;
}

axs_status axs_soft_slab_stepper_destroy(axs_soft_slab_stepper* st) {
  delete st;
  return AXS_OK;
}

axs_status axs_soft_slab_stepper_run(axs_soft_slab_stepper* st, int nsteps, int use_graph,
                                     void* stream) {
  AXS_CHECK_ARG(st, "null stepper");
  AXS_CHECK_ARG(nsteps >= 0, "negative step count");
  AXS_CHECK_ARG(!use_graph || stream, "graph capture needs a non-null stream");
  AXS_HIP(st->impl.run(nsteps, use_graph != 0, as_stream(stream)));
  return AXS_OK;
}

}  // extern "C"
