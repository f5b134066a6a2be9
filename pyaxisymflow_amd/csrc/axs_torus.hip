// Device-resident timestep of the dilating-torus driver
// (examples/TorusStreaming/torus_streaming.py:129-233, implicit diffusion as
// the driver sets it): a torus whose cross-section centre oscillates in r as
// a sum of modes, its surface velocity imposed by Brinkman penalisation and
// the dilation's source flow by a velocity potential.  One step, in the
// reference's order:
//   taper         kill_boundary_vorticity_sine_z then _r (:133-134), into a
//                 second buffer (the taper of a cell is a weight times one
//                 stored value; weights from the host)
//   dct2, tridiag, dct3
//                 psi (FastDiagonalisationStokesSolver.py:99-156)
//   velocity      compute_velocity_from_psi_unb (:137-138)
//   scalars       R_cm(t) = R_cm + e a sum A_k sin(w_k t) and the dilation
//                 rate e a sum A_k w_k cos(w_k t) (:168, :173-175)
//   mask          char_func of the moved torus (:169-170)
//   divergence    rhs = char_func * rate / r (:173-176)
//   potential, velocity from phi, u += grad phi
//                 (FastDiagonalisationPotentialSolver.py, compute_velocity_
//                 from_phi.py; :176-179)
//   penalise      u_pen = Brinkman(u) towards (0, rate) (:202-212), omega +=
//                 the interior curl of u_pen - u (:214-217)
//   advect        ENO3 with the unpenalised u (:220) -> home buffer
//   diffuse       implicit Euler diffusion at the fixed dt, in place (:223-224)
//   finalize      t += dt, it += 1
// FP64; every scalar on the device; steps captured into hipGraphs.
#include <new>
#include <vector>

#include "axs_eno3.hpp"
#include "axs_fused.hpp"
#include "axs_implicit.hpp"
#include "axs_potential.hpp"
#include "axs_stencil.hpp"
#include "axs_stokes.hpp"

namespace axs {

enum : int {
  kToT = 0,   // time
  kToDt,      // dt (fixed)
  kToDtByDx,  // dt / dx
  kToRcm,     // R_cm(t)
  kToRate,    // e a sum A_k w_k cos(w_k t)
  kToIt,      // step counter
  kToCount
};
static_assert(kToCount <= AXS_RIGID_SCALARS, "torus scalar slots exceed the array");

constexpr int kTorusMaxModes = 4;

struct TorusConfig {
  int nr, nz, kill_width, nmodes;
  double dx, nu, lam, moll_zone, a, AR2, e, Z_cm, R_cm, dt;
  double amp[kTorusMaxModes], omega[kTorusMaxModes];
};

struct TorusBuffers {
  double *vorticity, *psi, *u_z, *u_r, *u_z_pen, *u_r_pen, *char_func, *vel_phi, *scalars;
  const double *r, *z, *kill_wz, *kill_wr;
};

// ------------------------------------------------------------------ kernels

// kill_boundary_vorticity_sine_z then _r: the z ends of every row from
// columns w-1 and nz-w, then rows nr-w..nr-1 scaled copies of row nr-w,
// then row 0 = 0; the tapered value of (i, j) is a weight times the stored
// value at (min(i, nr-w), clamp(j, w-1, nz-w)).
__global__ void torus_taper_kernel(const double* __restrict__ w, double* __restrict__ out,
                                   const double* __restrict__ wz, const double* __restrict__ wr,
                                   int nr, int nz, int kw) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  const int si = (i >= nr - kw) ? nr - kw : i;
  const int sj = (j < kw) ? kw - 1 : ((j >= nz - kw) ? nz - kw : j);
  double t = w[(int64_t)si * nz + sj];
  if (j < kw)
    t = wz[j] * t;
  else if (j >= nz - kw)
    t = wz[kw + (j - (nz - kw))] * t;
  if (i >= nr - kw) t = wr[i - (nr - kw)] * t;
  if (i == 0) t = 0.0;
  out[(int64_t)i * nz + j] = t;
}

// One thread: the step's scalars (numpy's order: e a first, the mode sums
// left to right).
__global__ void torus_scalars_kernel(double* __restrict__ sc, TorusConfig cfg) {
  if (threadIdx.x != 0) return;
  const double t = sc[kToT];
  double ssum = 0.0, csum = 0.0;
  for (int k = 0; k < cfg.nmodes; ++k) {
    const double sk = cfg.amp[k] * sin(cfg.omega[k] * t);
    const double ck = (cfg.amp[k] * cfg.omega[k]) * cos(cfg.omega[k] * t);
    ssum = (k == 0) ? sk : ssum + sk;
    csum = (k == 0) ? ck : csum + ck;
  }
  const double ea = cfg.e * cfg.a;
  sc[kToRcm] = cfg.R_cm + ea * ssum;
  sc[kToRate] = ea * csum;
  sc[kToDt] = cfg.dt;
  sc[kToDtByDx] = cfg.dt / cfg.dx;
}

// char_func = smooth_Heaviside(a - sqrt((z - Z_cm)^2 + (r - R_cm(t))^2 / AR^2))
// and the potential's right-hand side char_func * (rate / r).
__global__ void torus_mask_kernel(double* __restrict__ chi, double* __restrict__ rhs,
                                  const double* __restrict__ r, const double* __restrict__ z,
                                  const double* __restrict__ sc, TorusConfig cfg) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= cfg.nz) return;
  const int64_t c = (int64_t)i * cfg.nz + j;
  const double dz = z[j] - cfg.Z_cm, dr = r[i] - sc[kToRcm];
  const double phi = cfg.a - sqrt(dz * dz + (dr * dr) / cfg.AR2);
  const double h = smooth_heaviside(phi, cfg.moll_zone);
  chi[c] = h;
  rhs[c] = h * (sc[kToRate] / r[i]);
}

__global__ void torus_add_kernel(double* __restrict__ uz, double* __restrict__ ur,
                                 const double* __restrict__ dz, const double* __restrict__ dr,
                                 int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  uz[k] = uz[k] + dz[k];
  ur[k] = ur[k] + dr[k];
}

// brinkmann_penalize(lam, dt, chi, 0, rate, u_z, u_r, u_z_pen, u_r_pen), and
// the correction u_pen - u whose curl is added to omega.
__global__ void torus_penalize_kernel(const double* __restrict__ uz, const double* __restrict__ ur,
                                      double* __restrict__ pz, double* __restrict__ pr,
                                      double* __restrict__ dz, double* __restrict__ dr,
                                      const double* __restrict__ chi,
                                      const double* __restrict__ sc, double lam, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double ldc = (lam * sc[kToDt]) * chi[k];
  const double den = 1.0 + ldc;
  const double gz = uz[k], gr = ur[k];
  const double a = (gz + ldc * 0.0) / den;
  const double b = (gr + ldc * sc[kToRate]) / den;
  pz[k] = a;
  pr[k] = b;
  dz[k] = a - gz;
  dr[k] = b - gr;
}

// omega += p, p the interior curl of (dz, dr) and 0 on the border.
__global__ void torus_add_curl_kernel(double* __restrict__ w, const double* __restrict__ dz,
                                      const double* __restrict__ dr, int nr, int nz, double dx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  const int64_t c = (int64_t)i * nz + j;
  double p = 0.0;
  if (i >= 1 && i <= nr - 2 && j >= 1 && j <= nz - 2) {
    const double two_dx = 2.0 * dx;
    p = (dr[c + 1] - dr[c - 1]) / two_dx - (dz[c + nz] - dz[c - nz]) / two_dx;
  }
  w[c] = w[c] + p;
}

__global__ void torus_finalize_kernel(double* __restrict__ sc) {
  if (threadIdx.x != 0) return;
  sc[kToT] = sc[kToT] + sc[kToDt];
  sc[kToIt] = sc[kToIt] + 1.0;
}

__global__ void torus_zero_kernel(double* __restrict__ p, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0.0;
}

// ------------------------------------------------------------------ stepper

struct TorusStepper {
  TorusConfig cfg{};
  TorusBuffers buf{};
  StokesPlan* plan = nullptr;
  PotentialPlan* pot = nullptr;
  ImplicitPlan* imp = nullptr;
  std::vector<void*> allocs;
  double *wt = nullptr, *rhs = nullptr, *gz = nullptr, *gr = nullptr, *dz = nullptr,
         *dr = nullptr;
  hipGraph_t graph = nullptr, chunk_graph = nullptr;
  hipGraphExec_t graph_exec = nullptr, chunk_exec = nullptr;
  hipStream_t graph_stream = nullptr;
  static constexpr int kChunkSteps = 16;

  hipError_t init(const TorusConfig& c, const TorusBuffers& b) {
    cfg = c;
    buf = b;
    plan = new (std::nothrow) StokesPlan();
    pot = new (std::nothrow) PotentialPlan();
    imp = new (std::nothrow) ImplicitPlan();
    if (!plan || !pot || !imp) return hipErrorOutOfMemory;
    hipError_t e = plan->init(cfg.nr, cfg.nz, cfg.dx);
    if (e == hipSuccess) e = pot->init(cfg.nr, cfg.nz, cfg.dx);
    if (e == hipSuccess) e = imp->init(cfg.nr, cfg.nz, cfg.dx, cfg.nu * cfg.dt);
    if (e != hipSuccess) return e;
    const size_t n = (size_t)cfg.nr * cfg.nz;
    for (double** f : {&wt, &rhs, &gz, &gr, &dz, &dr}) {
      void* q = nullptr;
      if ((e = hipMalloc(&q, sizeof(double) * n)) != hipSuccess) return e;
      allocs.push_back(q);
      *f = static_cast<double*>(q);
      torus_zero_kernel<<<ceil_div((int64_t)n, 256), 256>>>(*f, (int64_t)n);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
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

  ~TorusStepper() {
    drop_graphs();
    for (void* p : allocs) (void)hipFree(p);
    delete plan;
    delete pot;
    delete imp;
  }

  hipError_t enqueue_step(hipStream_t s) {
    const int nr = cfg.nr, nz = cfg.nz;
    const int64_t n = (int64_t)nr * nz;
    const dim3 g2(ceil_div(nz, 256), nr);
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
    torus_taper_kernel<<<g2, 256, 0, s>>>(buf.vorticity, wt, buf.kill_wz, buf.kill_wr, nr, nz,
                                          cfg.kill_width);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(launch_dct2_rows<double>(*plan, wt, plan->spec, plan->rvec, nr, s));
    AXS_TRY_STEP(launch_tridiag<double>(*plan, plan->spec, s));
    AXS_TRY_STEP(launch_dct3_rows<double>(*plan, plan->spec, buf.psi, nr, s));
    AXS_TRY_STEP(launch_velocity_from_psi<double>(buf.u_z, buf.u_r, buf.psi, buf.r, nr, nz,
                                                  cfg.dx, s));
    torus_scalars_kernel<<<1, 64, 0, s>>>(sc, cfg);
    AXS_TRY_LAUNCH();
    torus_mask_kernel<<<g2, 256, 0, s>>>(buf.char_func, rhs, buf.r, buf.z, sc, cfg);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(pot->solve(buf.vel_phi, rhs, s));
    AXS_TRY_STEP(launch_velocity_from_phi(gz, gr, buf.vel_phi, nr, nz, cfg.dx, s));
    torus_add_kernel<<<ceil_div(n, 256), 256, 0, s>>>(buf.u_z, buf.u_r, gz, gr, n);
    AXS_TRY_LAUNCH();
    torus_penalize_kernel<<<ceil_div(n, 256), 256, 0, s>>>(buf.u_z, buf.u_r, buf.u_z_pen,
                                                           buf.u_r_pen, dz, dr, buf.char_func, sc,
                                                           cfg.lam, n);
    AXS_TRY_LAUNCH();
    torus_add_curl_kernel<<<g2, 256, 0, s>>>(wt, dz, dr, nr, nz, cfg.dx);
    AXS_TRY_LAUNCH();
    AXS_TRY_STEP(launch_eno3_conservative<double>(wt, buf.u_z, buf.u_r, buf.vorticity, nr, nz,
                                                  Scalar<double>{0.0, sc + kToDtByDx}, -1.0, s));
    AXS_TRY_STEP(imp->launch(buf.vorticity, 1, s));
    torus_finalize_kernel<<<1, 64, 0, s>>>(sc);
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
      hipGraph_t* gr_ = q == 0 ? &graph : &chunk_graph;
      hipGraphExec_t* x = q == 0 ? &graph_exec : &chunk_exec;
      e = hipStreamEndCapture(s, gr_);
      if (le != hipSuccess) return le;
      if (e != hipSuccess) return e;
      if ((e = hipGraphInstantiate(x, *gr_, nullptr, nullptr, 0)) != hipSuccess) return e;
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

struct axs_torus_stepper {
  TorusStepper impl;
};

extern "C" {

axs_status axs_torus_stepper_create(const axs_torus_config* c, const axs_torus_buffers* b,
                                    axs_torus_stepper** out) {
  AXS_CHECK_ARG(c && b && out, "null pointer");
  *out = nullptr;
  AXS_CHECK_ARG(c->nr >= 4 && c->nr <= 4096 && c->nz >= 8 && stokes_nz_supported(c->nz),
                "grid not supported");
  AXS_CHECK_ARG(any_dft_length(c->nz + 1, nullptr) <= kAnyMaxL,
                "nz + 1 too long for the implicit diffusion's sine rows");
  AXS_CHECK_ARG(c->kill_width >= 2 && 2 * c->kill_width <= c->nz && c->kill_width < c->nr,
                "bad kill_width for the grid");
  AXS_CHECK_ARG(c->nmodes >= 1 && c->nmodes <= kTorusMaxModes, "nmodes must be in [1, 4]");
  AXS_CHECK_ARG(c->dx > 0 && c->nu > 0 && c->dt > 0 && c->lam >= 0 && c->moll_zone > 0 &&
                    c->a > 0 && c->AR > 0,
                "dx, nu, dt, moll_zone, a and AR must be positive, lam non-negative");
  AXS_CHECK_ARG(b->vorticity && b->psi && b->u_z && b->u_r && b->u_z_pen && b->u_r_pen &&
                    b->char_func && b->vel_phi && b->scalars && b->r && b->z && b->kill_wz &&
                    b->kill_wr,
                "null buffer");
  axs_torus_stepper* st = new (std::nothrow) axs_torus_stepper();
  if (!st) {
    set_error("out of host memory");
    return AXS_ERR_ALLOC;
  }
  TorusConfig cfg{};
  cfg.nr = c->nr;
  cfg.nz = c->nz;
  cfg.kill_width = c->kill_width;
  cfg.nmodes = c->nmodes;
  cfg.dx = c->dx;
  cfg.nu = c->nu;
  cfg.lam = c->lam;
  cfg.moll_zone = c->moll_zone;
  cfg.a = c->a;
  cfg.AR2 = c->AR * c->AR;  // AR**2
  cfg.e = c->e;
  cfg.Z_cm = c->Z_cm;
  cfg.R_cm = c->R_cm;
  cfg.dt = c->dt;
  for (int k = 0; k < kTorusMaxModes; ++k) {
    cfg.amp[k] = k < c->nmodes ? c->amp[k] : 0.0;
    cfg.omega[k] = k < c->nmodes ? c->omega[k] : 0.0;
  }
  TorusBuffers bb{b->vorticity, b->psi,     b->u_z,     b->u_r, b->u_z_pen,
                  b->u_r_pen,   b->char_func, b->vel_phi, b->scalars, b->r,
                  b->z,         b->kill_wz, b->kill_wr};
  hipError_t e = st->impl.init(cfg, bb);
  if (e != hipSuccess) {
    set_error(std::string("torus stepper: ") + hipGetErrorString(e));
    delete st;
    return AXS_ERR_HIP;
  }
  *out = st;
  return AXS_OK;
}

axs_status axs_torus_stepper_destroy(axs_torus_stepper* st) {
  delete st;
  return AXS_OK;
}

axs_status axs_torus_stepper_run(axs_torus_stepper* st, int nsteps, int use_graph, void* stream) {
  AXS_CHECK_ARG(st, "null stepper");
  AXS_CHECK_ARG(nsteps >= 0, "negative step count");
  AXS_CHECK_ARG(!use_graph || stream, "graph capture needs a non-null stream");
  AXS_HIP(st->impl.run(nsteps, use_graph != 0, as_stream(stream)));
  return AXS_OK;
}

}  // extern "C"
