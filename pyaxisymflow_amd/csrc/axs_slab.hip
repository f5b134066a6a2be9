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
struct SlabFields {
  double* f[3];
};
__global__ void slab_ghost_kernel(SlabFields fs, int nz, int g) {
  const int q = threadIdx.x;
  if (q >= 2 * g) return;
  double* row = fs.f[blockIdx.z] + (int64_t)blockIdx.x * nz;
  if (q < g)
    row[q] = row[nz - 2 * g + q];
  else
    row[nz - g + (q - g)] = row[q];
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
    slab_ghost_kernel<<<dim3(cfg.nr, 1, nf), 64, 0, s>>>(SlabFields{{a, b, c}}, cfg.nz, cfg.ghost);
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

}  // namespace axs

using namespace axs;

struct axs_slab_stepper {
  SlabStepper impl;
};

extern "C" {

axs_status axs_slab_stepper_create(const axs_slab_config* c, const axs_slab_buffers* b,
                                   axs_slab_stepper** out) {
  AXS_CHECK_ARG(c && b && out, "null pointer");
  *out = nullptr;
  AXS_CHECK_ARG(c->ghost >= 1 && c->nz - 2 * c->ghost >= 2 && c->nz >= 4 * c->ghost,
                "need nz >= 4 ghost and at least 2 interior columns");
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

}  // extern "C"
