// Device-resident timestep of the rigid-body drivers (row A16):
// examples/FlowPastSphere/flow_past_sphere.py:107-188 (static body, free
// stream with sine ramp, drag coefficient) and
// examples/SpheroidStreaming/spheroid_streaming.py:129-246 (oscillating
// spheroid in quiescent fluid: mask rebuilt every step, penalisation
// velocity U_0 cos(omega t), dt clipped to the period and to tEnd, psi
// averaged over each period).
//
// One step, in the reference's order, as six passes (a stage is skipped
// when its feature is off):
//   taper_dct2_rows   kill taper of the stored field on the fly, DCT-II of
//                     r * omega along z; one extra workgroup per case
//                     completes the previous step (force record, t, it)
//   tridiag_columns   one radial tridiagonal solve per z-mode
//   dct3_rows         inverse DCT -> psi
//   velocity_max      u from psi (+ free stream), max|u| (atomicMax slots)
//   body_mask         chi = H(phi(t)) of the moving body
//   step_fused        dt rule, u from psi, Brinkman, taper + curl(u_pen - u),
//                     force partials, psi average, ENO3, both RK2 stages ->
//                     the other buffer of the vorticity pair
// The stored vorticity alternates between the home buffer and work0; one
// graph is captured per direction.  A run ends with a standalone completion
// of its last step and, if it ended on work0, a copy home, so the home
// buffer, t, it and the force history are current whenever a run returns.
// Every scalar (t, dt, ramp factor, penalisation velocity, period timer,
// force history) stays in device memory, so a step is captured once into a
// hipGraph and replayed.  No memset nodes are recorded (fills are kernels).
//
// Precision: T = double is the plain FP64 path.  T = float stores every
// field in FP32 and runs the stencils and z-transforms in FP32, while the
// radial tridiagonal elimination, the mask level set, and all reductions
// and step scalars stay FP64.
#include "axs_fused.hpp"
#include "axs_implicit.hpp"
#include "axs_stepper.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <new>
#include <vector>

namespace axs {

// ------------------------------------------------------------------ kernels

// chi = smooth_Heaviside(phi) of the spheroid centred at z = sc[BodyZ]
// (spheroid_streaming.py:207-212), level set and blend in FP64.
template <typename T>
__global__ void body_mask_kernel(T* __restrict__ chi, const double* __restrict__ r,
                                 const double* __restrict__ z, int nr, int nz,
                                 const double* __restrict__ sc,
                                 const RigidConfig* __restrict__ cfgs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  const int cb = blockIdx.z;  // case of a batch
  const RigidConfig& cfg = cfgs[cb];
  chi += (int64_t)cb * nr * nz;
  sc += (int64_t)cb * AXS_RIGID_SCALARS;
  const double zc = sc[kScBodyZ];
  const double dz = z[j] - zc, dr = r[i] - cfg.body_r0;
  const double phi = cfg.body_radius - sqrt(dz * dz / cfg.body_aspect2 + dr * dr);
  const double w = cfg.moll_zone;
  double h = (phi >= w) ? 1.0 : 0.0;
  if (fabs(phi) < w) h = h + 0.5 * (1.0 + phi / w + sin(M_PI * phi / w) / M_PI);
  chi[(int64_t)i * nz + j] = T(h);
}

// Step scalars for the time now in sc[T] (without advancing); clears the
// velocity maximum slots.
// One block per case.
__global__ void init_scalars_kernel(double* __restrict__ sc, double* __restrict__ slots,
                                    const RigidConfig* __restrict__ cfgs) {
  const int cb = blockIdx.x;
  sc += (int64_t)cb * AXS_RIGID_SCALARS;
  slots += (int64_t)cb * kSlotsPerCase;
  if (threadIdx.x < kMaxSlots) slots[threadIdx.x * kSlotStride] = 0.0;
  if (threadIdx.x == 0) {
    sc[kScPending] = 0.0;
    next_step_scalars(sc, cfgs[cb], sc[kScT]);
  }
}

// ------------------------------------------------------------------ stepper

const char* const kRigidStageNames[kRigidStages] = {
    "taper_dct2_rows", "tridiag_columns", "dct3_rows",         "velocity_max",
    "body_mask",       "step_fused",      "implicit_diffusion"};

template <typename T>
__global__ void copy_field_kernel(T* __restrict__ dst, const T* __restrict__ src, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[k];
}

void RigidStepperBase::drop_graphs() {
  for (int k = 0; k < kGraphSizes; ++k)
    for (int p = 0; p < 2; ++p)
      for (int end = 0; end < 2; ++end) {
        if (graph_execs[k][p][end]) (void)hipGraphExecDestroy(graph_execs[k][p][end]);
        if (graphs[k][p][end]) (void)hipGraphDestroy(graphs[k][p][end]);
        graph_execs[k][p][end] = nullptr;
        graphs[k][p][end] = nullptr;
      }
  graph_stream = nullptr;
}

RigidStepperBase::~RigidStepperBase() {
  drop_graphs();
  if (partials) (void)hipFree(partials);
  if (slots) (void)hipFree(slots);
  if (spec) (void)hipFree(spec);
  if (dcfg) (void)hipFree(dcfg);
  delete plan;
  delete implicit;
}

hipError_t RigidStepperBase::init(const RigidConfig* cfgs, int nbatch, const RigidBuffers& b) {
  cfg = cfgs[0];
  batch = nbatch;
  buf = b;
  plan = new (std::nothrow) StokesPlan();
  if (!plan) return hipErrorOutOfMemory;
  // the plan picks the half-length transforms or the any-length ones
  hipError_t e = plan->init(cfg.nr, cfg.nz, cfg.dx);
  if (e != hipSuccess) return e;
  if (cfg.implicit_diffusion) {
    // implicit_diffusion_solver.py:16: nu_times_dt = time_step * kinematic_viscosity
    implicit = new (std::nothrow) ImplicitPlan();
    if (!implicit) return hipErrorOutOfMemory;
    e = implicit->init(cfg.nr, cfg.nz, cfg.dx, cfg.dt_diffusive * cfg.nu, nbatch);
    if (e != hipSuccess) return e;
  }
  blocks_x = ceil_div(cfg.nz, 256);
  npartials = fused_blocks(cfg.nr, cfg.nz);
  const size_t cells = (size_t)cfg.nr * cfg.nz * batch;
  e = hipMalloc(reinterpret_cast<void**>(&partials), sizeof(double) * (size_t)npartials * batch);
  if (e != hipSuccess) return e;
  e = hipMalloc(reinterpret_cast<void**>(&slots), sizeof(double) * kSlotsPerCase * batch);
  if (e != hipSuccess) return e;
  e = hipMalloc(&spec, sizeof(double) * cells);
  if (e != hipSuccess) return e;
  e = hipMalloc(reinterpret_cast<void**>(&dcfg), sizeof(RigidConfig) * batch);
  if (e != hipSuccess) return e;
  e = hipMemcpy(dcfg, cfgs, sizeof(RigidConfig) * batch, hipMemcpyHostToDevice);
  if (e != hipSuccess) return e;
  init_scalars_kernel<<<batch, 64>>>(buf.scalars, slots, dcfg);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipDeviceSynchronize();
}

bool RigidStepperBase::stage_enabled(int stage) const {
  if (stage == 4) return cfg.moving_body != 0;
  if (stage == 6) return cfg.implicit_diffusion != 0;
  return stage >= 0 && stage < kRigidStages;
}

template <typename T>
hipError_t RigidStepper<T>::enqueue_stage(int stage, hipStream_t s) {
  const int nr = cfg.nr, nz = cfg.nz, w = cfg.kill_width;
  double* sc = buf.scalars;
  T* home = static_cast<T*>(buf.vorticity);
  T* other = static_cast<T*>(buf.work0);
  T* cur = parity ? other : home;
  T* nxt = parity ? home : other;
  T* psi = static_cast<T*>(buf.psi);
  T* chi = static_cast<T*>(buf.char_func);
  T* sp = static_cast<T*>(spec);
  const int B = batch;
  if (!stage_enabled(stage)) return hipSuccess;
  switch (stage) {
    case 0: {
      // the previous step completes in an extra workgroup of this pass
      StepFinalize fin;
      fin.partials = partials;
      fin.count = npartials;
      fin.slots = slots;
      fin.sc = sc;
      fin.drag = buf.drag;
      fin.cfgs = dcfg;
      return launch_taper_dct2_rows<T>(*plan, cur, nullptr, sp, buf.kill_wz, buf.kill_wr, w, s, B,
                                       &fin);
    }
    case 1:
      return launch_tridiag<T>(*plan, sp, s, B);
    case 2:
      return launch_dct3_rows<T>(*plan, sp, psi, nr, s, B);
    case 3:
      return launch_velocity_max<T>(psi, buf.r, nr, nz, T(cfg.dx), sc, slots, dcfg, B, s);
    case 4:
      body_mask_kernel<T><<<dim3(blocks_x, nr, B), 256, 0, s>>>(chi, buf.r, buf.z, nr, nz, sc,
                                                                 dcfg);
      return hipGetLastError();
    case 5:
      return launch_step_fused<T>(psi, chi, cur, nxt,
                                  cfg.average_psi ? static_cast<T*>(buf.avg_psi) : nullptr,
                                  buf.r, buf.kill_wz, buf.kill_wr, w, nr, nz, T(cfg.dx), sc, slots,
                                  partials, dcfg, B, s, cfg.implicit_diffusion == 0);
    case 6:
      // implicit Euler diffusion of the step's new vorticity, in place
      // (flow_past_sphere.py:179-181); FP64 only (checked at creation)
      if constexpr (sizeof(T) == 8)
        return implicit->launch(reinterpret_cast<double*>(nxt), B, s);
      else
        return hipErrorInvalidValue;
    default:
      return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t RigidStepper<T>::settle(hipStream_t s) {
  // complete the last step (a no-op when nothing is pending)
  hipError_t e = launch_finalize_step(partials, npartials, slots, buf.scalars, buf.drag, dcfg,
                                      batch, s);
  if (e != hipSuccess) return e;
  if (!parity) return hipSuccess;
  const int64_t n = (int64_t)cfg.nr * cfg.nz * batch;
  copy_field_kernel<T><<<ceil_div(n, 256), 256, 0, s>>>(static_cast<T*>(buf.vorticity),
                                                        static_cast<const T*>(buf.work0), n);
  parity = 0;
  return hipGetLastError();
}

template <typename T>
hipError_t RigidStepper<T>::velocity_out(hipStream_t s) {
  return launch_velocity_out<T>(static_cast<const T*>(buf.psi),
                                static_cast<const T*>(buf.char_func),
                                static_cast<T*>(buf.u_pen_z), static_cast<T*>(buf.u_pen_r), buf.r,
                                cfg.nr, cfg.nz, T(cfg.dx), buf.scalars, dcfg, batch, s);
}

hipError_t RigidStepperBase::enqueue_step(hipStream_t s) {
  for (int q = 0; q < kRigidStages; ++q) {
    hipError_t e = enqueue_stage(q, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t RigidStepperBase::run(int nsteps, bool use_graph, hipStream_t s) {
  if (!use_graph) {
    for (int q = 0; q < nsteps; ++q) {
      hipError_t e = enqueue_step(s);
      if (e != hipSuccess) return e;
      parity ^= 1;
    }
    return settle(s);
  }
  if (graph_stream != s) {
    hipError_t e = capture_graphs(s);
    if (e != hipSuccess) return e;
  }
  if (nsteps == 0) return settle(s);
  int left = nsteps;
  while (left > 0) {
    // the largest graph that fits: 16 steps, else the top bit of what is left
    const int k = left >= kChunkSteps ? kGraphSizes - 1 : 31 - __builtin_clz((unsigned)left);
    const int n = 1 << k;
    const int end = left == n ? 1 : 0;
    hipError_t e = hipGraphLaunch(graph_execs[k][parity][end], s);
    if (e != hipSuccess) return e;
    ++(n == kChunkSteps ? chunk_launches : step_launches);
    left -= n;
    parity = end ? 0 : (parity ^ (n & 1));
  }
  return hipSuccess;
}

hipError_t RigidStepperBase::capture_graphs(hipStream_t s) {
  drop_graphs();
  const int saved = parity;
  for (int k = 0; k < kGraphSizes; ++k)
    for (int p = 0; p < 2; ++p)
      for (int end = 0; end < 2; ++end) {
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
          parity = saved;
          return e;
        }
        parity = p;
        hipError_t le = hipSuccess;
        for (int q = 0; q < (1 << k) && le == hipSuccess; ++q) {
          le = enqueue_step(s);
          parity ^= 1;
        }
        if (le == hipSuccess && end) le = settle(s);
        e = hipStreamEndCapture(s, &graphs[k][p][end]);
        if (le == hipSuccess && e == hipSuccess)
          e = hipGraphInstantiate(&graph_execs[k][p][end], graphs[k][p][end], nullptr, nullptr, 0);
        if (le != hipSuccess || e != hipSuccess) {
          parity = saved;
          return le != hipSuccess ? le : e;
        }
      }
  parity = saved;
  graph_stream = s;
  return hipSuccess;
}

// ------------------------------------------------------------ uneven batches

static bool same_grid(const RigidConfig& a, const RigidConfig& b) {
  return a.nr == b.nr && a.nz == b.nz && a.dx == b.dx;
}

// Radix plan of the half-length transform (radices 4, 2, 3, 5); false if
// nz / 2 has another prime factor.
static bool fft_args(int nz, FftPlanArgs& fp) {
  fp = FftPlanArgs{};
  fp.M = nz / 2;
  int m = fp.M;
  for (int rdx : {4, 2, 3, 5})
    while (m % rdx == 0 && fp.nstages < kMaxFftStages) {
      fp.radix[fp.nstages++] = rdx;
      m /= rdx;
    }
  return m == 1;
}

// Lays out the velocity_max tiles of every group for tiles of tr rows;
// returns the tiles of the whole batch.
static int set_vmax_tiles(std::vector<RaggedGroup>& groups, int tr) {
  int vt = 0;
  for (RaggedGroup& G : groups) {
    G.vtiles = vmax_tiles(G.nr, G.nz, tr);
    G.vt0 = vt;
    vt += G.ncase * G.vtiles;
  }
  return vt;
}

// Groups (runs of one grid) and their plans for radial segments of L rows.
static hipError_t ragged_groups(const RigidConfig* cfgs, int nbatch, const RigidBuffers& b, int L,
                                std::vector<StokesPlan*>& plans, std::vector<RaggedGroup>& groups,
                                int& tri_blocks, size_t& tri_shm) {
  const int kw = cfgs[0].kill_width;
  int64_t cell = 0, r_off = 0, w_off = 0;
  int row = 0, wrow = 0, tri = 0, vt = 0, ft = 0;
  tri_shm = 0;
  for (int q = 0; q < nbatch;) {
    int e = q;
    while (e < nbatch && same_grid(cfgs[e], cfgs[q])) ++e;
    if ((int)groups.size() == kMaxRaggedGroups) return hipErrorInvalidValue;
    const int nr = cfgs[q].nr, nz = cfgs[q].nz, nc = e - q;
    StokesPlan* pl = new (std::nothrow) StokesPlan();
    if (!pl) return hipErrorOutOfMemory;
    plans.push_back(pl);
    if (!fft_args(nz, pl->fft)) return hipErrorInvalidValue;
    pl->force_L = L;
    hipError_t err = pl->init(nr, nz, cfgs[q].dx);
    if (err != hipSuccess) return err;
    if (pl->tri_G) return hipErrorInvalidValue;  // coefficients must fit the LDS
    RaggedGroup G{};
    G.nr = nr;
    G.nz = nz;
    G.ncase = nc;
    G.case0 = q;
    // rows of a 256-thread workgroup: bs threads per row cover every stage
    // of the row transform (bs >= M / 2)
    G.bs = 64;
    while (G.bs < nz / 4 && G.bs < 256) G.bs *= 2;
    G.rpb = 256 / G.bs;
    while (G.rpb > 1 && nr % G.rpb) {
      G.rpb /= 2;
      G.bs *= 2;
    }
    G.rb0 = row;
    // half-lengths with a one-wave register FFT: the one-grid kernels run
    // the group's row passes over its cases
    G.wave = wave_rows(nz / 2) ? 1 : 0;
    G.wr0 = wrow;
    G.K = pl->tri_K;
    G.P = pl->tri_P;
    G.tchunks = (int)ceil_div(nz, G.K);
    G.tri0 = tri;
    G.vz = vmax_tiles_z(nz);
    G.vtiles = vmax_tiles(nr, nz, 16);  // set_vmax_tiles re-lays them out
    G.vt0 = vt;
    G.fz = fused_tiles_z(nz);
    G.ftiles = fused_blocks(nr, nz);
    G.ft0 = ft;
    G.part0 = ft;
    G.cell0 = cell;
    G.dx = cfgs[q].dx;
    G.fp = pl->fft;
    G.tw = pl->tw;
    G.e1 = pl->e1;
    G.e2 = pl->e2;
    G.rvec = pl->rvec;
    G.a = pl->a;
    G.b = pl->b;
    G.c = pl->c;
    G.lam = pl->lam;
    G.cseg = pl->cseg;
    G.r = b.r + r_off;  // the group's first case's tables
    G.kill_wz = b.kill_wz + 2 * w_off;
    G.kill_wr = b.kill_wr + w_off;
    if (G.wave)
      wrow += nc * nr;
    else
      row += nc * nr / G.rpb;
    tri += nc * G.tchunks;
    vt += nc * G.vtiles;
    ft += nc * G.ftiles;
    cell += (int64_t)nc * nr * nz;
    r_off += (int64_t)nc * nr;
    w_off += (int64_t)nc * kw;
    tri_shm = std::max(tri_shm, tri_lds_bytes(nr, L, G.K, G.P, false));
    groups.push_back(G);
    q = e;
  }
  tri_blocks = tri;
  return hipSuccess;
}

static void free_plans(std::vector<StokesPlan*>& plans) {
  for (StokesPlan* p : plans) delete p;
  plans.clear();
}

template <typename T>
hipError_t RaggedStepper<T>::init_ragged(const RigidConfig* cfgs, int nbatch,
                                         const RigidBuffers& b) {
  cfg = cfgs[0];
  batch = nbatch;
  buf = b;
  cells = 0;
  for (int q = 0; q < nbatch; ++q) cells += (int64_t)cfgs[q].nr * cfgs[q].nz;
  hipError_t e = hipMalloc(&spec, sizeof(double) * (size_t)cells);
  if (e != hipSuccess) return e;
  // Radial segment length: AXS_TRIDIAG_LK pins it (its K is ignored: an
  // uneven launch fills 1024 threads per group), AXS_NO_TUNE takes 8, else
  // the whole batch's radial pass is timed for each candidate.
  std::vector<int> cands{8, 4, 16};
  if (const char* pin = std::getenv("AXS_TRIDIAG_LK")) {
    int L = 0;
    if (std::sscanf(pin, "%d", &L) != 1 || (L != 4 && L != 8 && L != 16))
      return hipErrorInvalidValue;
    cands = {L};
  } else if (std::getenv("AXS_NO_TUNE")) {
    cands = {8};
  }
  float best = 1e30f;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (cands.size() > 1) {
    if ((e = hipMemset(spec, 0, sizeof(double) * (size_t)cells)) != hipSuccess) return e;
    if ((e = hipEventCreate(&t0)) != hipSuccess) return e;
    if ((e = hipEventCreate(&t1)) != hipSuccess) return e;
  }
  for (int L : cands) {
    std::vector<StokesPlan*> pl;
    std::vector<RaggedGroup> gr;
    int tri = 0;
    size_t shm = 0;
    e = ragged_groups(cfgs, nbatch, b, L, pl, gr, tri, shm);
    if (e == hipErrorInvalidValue && cands.size() > 1) {  // L does not fit this batch
      free_plans(pl);
      continue;
    }
    if (e != hipSuccess) {
      free_plans(pl);
      break;
    }
    float ms = 0.f;
    if (cands.size() > 1) {
      RaggedGroup* dg = nullptr;
      e = hipMalloc(reinterpret_cast<void**>(&dg), sizeof(RaggedGroup) * gr.size());
      if (e == hipSuccess)
        e = hipMemcpy(dg, gr.data(), sizeof(RaggedGroup) * gr.size(), hipMemcpyHostToDevice);
      T* sp = static_cast<T*>(spec);
      for (int rep = 0; rep < 2 && e == hipSuccess; ++rep)
        e = launch_tridiag_ragged<T>(sp, dg, (int)gr.size(), tri, shm, L, 0);
      ms = 1e30f;
      for (int trial = 0; trial < 3 && e == hipSuccess; ++trial) {
        float t = 0.f;
        e = hipEventRecord(t0, 0);
        for (int rep = 0; rep < 5 && e == hipSuccess; ++rep)
          e = launch_tridiag_ragged<T>(sp, dg, (int)gr.size(), tri, shm, L, 0);
        if (e == hipSuccess) e = hipEventRecord(t1, 0);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, t0, t1);
        if (e == hipSuccess && t < ms) ms = t;
      }
      if (dg) (void)hipFree(dg);
      if (e != hipSuccess) {
        free_plans(pl);
        break;
      }
    }
    if (groups.empty() || ms < best) {
      best = ms;
      free_plans(plans);
      plans.swap(pl);
      groups.swap(gr);
      tri_blocks = tri;
      tri_shm = shm;
      tri_L = L;
    } else {
      free_plans(pl);
    }
  }
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  if (e != hipSuccess && groups.empty()) return e;
  if (groups.empty()) return hipErrorInvalidValue;
  rows = wave_blocks = vt_tiles = ft_tiles = maxM = 0;
  for (const RaggedGroup& G : groups) {
    vt_tiles += G.ncase * G.vtiles;
    ft_tiles += G.ncase * G.ftiles;
    if (G.wave) {  // one-wave register rows, one launch over all such groups
      wave_blocks += G.ncase * G.nr;
      continue;
    }
    rows += G.ncase * G.nr / G.rpb;  // row-pass workgroups
    // LDS complex entries per workgroup: rpb rows and the twiddle table
    maxM = std::max(maxM, (G.rpb + 1) * (G.nz / 2));
  }
  npartials = ft_tiles;
  e = hipMalloc(reinterpret_cast<void**>(&partials), sizeof(double) * (size_t)ft_tiles);
  if (e != hipSuccess) return e;
  e = hipMalloc(reinterpret_cast<void**>(&slots), sizeof(double) * kSlotsPerCase * batch);
  if (e != hipSuccess) return e;
  e = hipMalloc(reinterpret_cast<void**>(&dcfg), sizeof(RigidConfig) * batch);
  if (e != hipSuccess) return e;
  e = hipMemcpy(dcfg, cfgs, sizeof(RigidConfig) * batch, hipMemcpyHostToDevice);
  if (e != hipSuccess) return e;
  e = hipMalloc(reinterpret_cast<void**>(&dgroups), sizeof(RaggedGroup) * groups.size());
  if (e != hipSuccess) return e;
  // velocity_max tile height: AXS_VMAX_TR pins it, AXS_NO_TUNE takes 16,
  // else the whole batch's pass is timed for 16, 8 and 32 rows (before the
  // slots it writes are cleared below)
  std::vector<int> trs{16, 8, 32};
  if (const char* pin = std::getenv("AXS_VMAX_TR")) {
    int tr = 0;
    if (std::sscanf(pin, "%d", &tr) != 1 || (tr != 8 && tr != 16 && tr != 32))
      return hipErrorInvalidValue;
    trs = {tr};
  } else if (std::getenv("AXS_NO_TUNE")) {
    trs = {16};
  }
  vmax_tr = trs[0];
  if (trs.size() > 1) {
    if ((e = hipEventCreate(&t0)) != hipSuccess) return e;
    if ((e = hipEventCreate(&t1)) != hipSuccess) return e;
    const T* psi = static_cast<const T*>(buf.psi);
    float best_v = 1e30f;
    for (int tr : trs) {
      const int tiles = set_vmax_tiles(groups, tr);
      e = hipMemcpy(dgroups, groups.data(), sizeof(RaggedGroup) * groups.size(),
                    hipMemcpyHostToDevice);
      for (int rep = 0; rep < 2 && e == hipSuccess; ++rep)
        e = launch_velocity_max_ragged<T>(psi, buf.scalars, slots, dcfg, dgroups,
                                          (int)groups.size(), tiles, tr, 0);
      float ms = 1e30f;
      for (int trial = 0; trial < 3 && e == hipSuccess; ++trial) {
        float t = 0.f;
        e = hipEventRecord(t0, 0);
        for (int rep = 0; rep < 5 && e == hipSuccess; ++rep)
          e = launch_velocity_max_ragged<T>(psi, buf.scalars, slots, dcfg, dgroups,
                                            (int)groups.size(), tiles, tr, 0);
        if (e == hipSuccess) e = hipEventRecord(t1, 0);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, t0, t1);
        if (e == hipSuccess && t < ms) ms = t;
      }
      if (e != hipSuccess) break;
      if (ms < best_v) {
        best_v = ms;
        vmax_tr = tr;
      }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (e != hipSuccess) return e;
  }
  vt_tiles = set_vmax_tiles(groups, vmax_tr);
  if (std::getenv("AXS_TUNE_LOG"))
    std::fprintf(stderr, "uneven batch: radial segments L = %d, velocity_max tiles %d x 64\n",
                 tri_L, vmax_tr);
  e = hipMemcpy(dgroups, groups.data(), sizeof(RaggedGroup) * groups.size(),
                hipMemcpyHostToDevice);
  if (e != hipSuccess) return e;
  init_scalars_kernel<<<batch, 64>>>(buf.scalars, slots, dcfg);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipDeviceSynchronize();
}

template <typename T>
RaggedStepper<T>::~RaggedStepper() {
  free_plans(plans);
  if (dgroups) (void)hipFree(dgroups);
}

template <typename T>
hipError_t RaggedStepper<T>::enqueue_stage(int stage, hipStream_t s) {
  const int w = cfg.kill_width, ng = (int)groups.size();
  double* sc = buf.scalars;
  T* home = static_cast<T*>(buf.vorticity);
  T* other = static_cast<T*>(buf.work0);
  T* cur = parity ? other : home;
  T* nxt = parity ? home : other;
  T* psi = static_cast<T*>(buf.psi);
  T* chi = static_cast<T*>(buf.char_func);
  T* sp = static_cast<T*>(spec);
  if (!stage_enabled(stage)) return hipSuccess;
  switch (stage) {
    case 0: {
      StepFinalize fin;
      fin.partials = partials;
      fin.slots = slots;
      fin.sc = sc;
      fin.drag = buf.drag;
      fin.cfgs = dcfg;
      hipError_t e = launch_taper_dct2_ragged<T>(cur, sp, dgroups, ng, batch, rows, w, maxM, &fin,
                                                 s);
      if (e != hipSuccess) return e;
      return launch_wave_rows_ragged<T>(cur, sp, dgroups, ng, wave_blocks, w, true, s);
    }
    case 1:
      return launch_tridiag_ragged<T>(sp, dgroups, ng, tri_blocks, tri_shm, tri_L, s);
    case 2: {
      hipError_t e = rows > 0 ? launch_dct3_ragged<T>(sp, psi, dgroups, ng, rows, maxM, s)
                              : hipSuccess;
      if (e != hipSuccess) return e;
      return launch_wave_rows_ragged<T>(sp, psi, dgroups, ng, wave_blocks, w, false, s);
    }
    case 3:
      return launch_velocity_max_ragged<T>(psi, sc, slots, dcfg, dgroups, ng, vt_tiles, vmax_tr,
                                           s);
    case 4:
      return hipSuccess;  // no moving body in uneven batches
    case 5:
      return launch_step_fused_ragged<T>(psi, chi, cur, nxt,
                                         cfg.average_psi ? static_cast<T*>(buf.avg_psi) : nullptr,
                                         w, sc, slots, partials, dcfg, dgroups, ng, ft_tiles, s);
    default:
      return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t RaggedStepper<T>::settle(hipStream_t s) {
  hipError_t e = launch_finalize_ragged(partials, slots, buf.scalars, buf.drag, dcfg, dgroups,
                                        (int)groups.size(), batch, s);
  if (e != hipSuccess) return e;
  if (!parity) return hipSuccess;
  copy_field_kernel<T><<<ceil_div(cells, 256), 256, 0, s>>>(static_cast<T*>(buf.vorticity),
                                                            static_cast<const T*>(buf.work0),
                                                            cells);
  parity = 0;
  return hipGetLastError();
}

template <typename T>
hipError_t RaggedStepper<T>::velocity_out(hipStream_t s) {
  for (const RaggedGroup& G : groups) {
    hipError_t e = launch_velocity_out<T>(
        static_cast<const T*>(buf.psi) + G.cell0, static_cast<const T*>(buf.char_func) + G.cell0,
        static_cast<T*>(buf.u_pen_z) + G.cell0, static_cast<T*>(buf.u_pen_r) + G.cell0, G.r, G.nr,
        G.nz, T(G.dx), buf.scalars + (int64_t)G.case0 * AXS_RIGID_SCALARS, dcfg + G.case0,
        G.ncase, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template struct RigidStepper<double>;
template struct RigidStepper<float>;
template struct RaggedStepper<double>;
template struct RaggedStepper<float>;

}  // namespace axs

using namespace axs;

struct axs_rigid_stepper {
  RigidStepperBase* impl = nullptr;
  ~axs_rigid_stepper() { delete impl; }
};

extern "C" {

static RigidConfig to_internal(const axs_rigid_config& x) {
  RigidConfig c{};
  c.nr = x.nr;
  c.nz = x.nz;
  c.kill_width = x.kill_width;
  c.history = x.history;
  c.precision = x.precision;
  c.moving_body = x.moving_body;
  c.average_psi = x.average_psi;
  c.implicit_diffusion = x.implicit_diffusion;
  c.dx = x.dx;
  c.nu = x.nu;
  c.brink_lam = x.brink_lam;
  c.U0 = x.U0;
  c.T_ramp = x.T_ramp;
  c.CFL = x.CFL;
  c.eps = x.eps;
  c.dt_diffusive = x.dt_diffusive;
  c.osc_amp = x.osc_amp;
  c.osc_omega = x.osc_omega;
  c.drag_scale = x.drag_scale;
  c.drag_div = x.drag_div;
  c.dt_cap = x.dt_cap;
  c.period = x.period;
  c.t_end = x.t_end;
  c.body_z0 = x.body_z0;
  c.body_r0 = x.body_r0;
  c.body_aspect2 = x.body_aspect2;
  c.body_radius = x.body_radius;
  c.body_disp = x.body_disp;
  c.moll_zone = x.moll_zone;
  return c;
}

axs_status axs_rigid_batch_create(const axs_rigid_config* cfgs, int batch,
                                  const axs_rigid_buffers* bufs, axs_rigid_stepper** out) {
  AXS_CHECK_ARG(cfgs && bufs && out, "null pointer");
  *out = nullptr;
  AXS_CHECK_ARG(batch >= 1 && batch <= 4096, "batch must be in [1, 4096]");
  const axs_rigid_config& c0 = cfgs[0];
  AXS_CHECK_ARG(c0.nr >= 4 && c0.nz >= 8, "need nr >= 4 and nz >= 8");
  AXS_CHECK_ARG(c0.nr <= 4096 && stokes_nz_supported(c0.nz),
                "grid too large for the row/column passes");
  AXS_CHECK_ARG(c0.kill_width >= 2 && 2 * c0.kill_width <= c0.nz && c0.kill_width < c0.nr,
                "bad kill_width for the grid");
  AXS_CHECK_ARG(c0.precision == AXS_F64 || c0.precision == AXS_F32, "bad precision");
  AXS_CHECK_ARG(c0.history >= 0, "negative history");
  bool uneven = false;
  int runs = 1;
  for (int q = 0; q < batch; ++q) {
    const axs_rigid_config& c = cfgs[q];
    AXS_CHECK_ARG(c.kill_width == c0.kill_width && c.history == c0.history &&
                      c.precision == c0.precision && c.moving_body == c0.moving_body &&
                      c.average_psi == c0.average_psi,
                  "cases of a batch must share kill width, history, precision and flags");
    AXS_CHECK_ARG(c.nr >= 4 && c.nz >= 8, "need nr >= 4 and nz >= 8");
    AXS_CHECK_ARG(2 * c.kill_width <= c.nz && c.kill_width < c.nr, "bad kill_width for the grid");
    AXS_CHECK_ARG(c.dx > 0 && c.nu > 0 && c.CFL > 0, "dx, nu and CFL must be positive");
    AXS_CHECK_ARG(!c.moving_body || (c.body_aspect2 > 0 && c.moll_zone > 0),
                  "moving body needs aspect^2 > 0 and moll_zone > 0");
    if (c.nr != c0.nr || c.nz != c0.nz || c.dx != c0.dx) uneven = true;
    if (q > 0 && (c.nr != cfgs[q - 1].nr || c.nz != cfgs[q - 1].nz || c.dx != cfgs[q - 1].dx))
      ++runs;
  }
  if (uneven) {
    // one launch per pass over cases of several grids (axs_ragged.hpp)
    AXS_CHECK_ARG(!c0.moving_body, "a batch of uneven grids takes no moving body");
    AXS_CHECK_ARG(runs <= kMaxRaggedGroups,
                  "a batch of uneven grids takes at most 16 runs of one grid (order the cases)");
    for (int q = 0; q < batch; ++q)
      AXS_CHECK_ARG(cfgs[q].nz <= 1024 && cfgs[q].nr <= 4096,
                    "a batch of uneven grids takes nz <= 1024 and nr <= 4096");
  }
  if (c0.implicit_diffusion) {
    AXS_CHECK_ARG(!uneven && c0.precision == AXS_F64,
                  "implicit diffusion runs in the FP64 one-grid stepper");
    for (int q = 0; q < batch; ++q)
      AXS_CHECK_ARG(cfgs[q].implicit_diffusion &&
                        cfgs[q].dt_diffusive * cfgs[q].nu == c0.dt_diffusive * c0.nu &&
                        cfgs[q].dt_diffusive > 0.0,
                    "implicit diffusion: every case of a batch needs the same nu * dt > 0");
  }
  AXS_CHECK_ARG(bufs->vorticity && bufs->psi && bufs->u_pen_z && bufs->u_pen_r &&
                    bufs->char_func && bufs->work0 && bufs->scalars && bufs->r && bufs->kill_wz &&
                    bufs->kill_wr,
                "null buffer");
  AXS_CHECK_ARG(!c0.average_psi || bufs->avg_psi, "average_psi needs the avg_psi buffer");
  AXS_CHECK_ARG(!c0.moving_body || bufs->z, "moving body needs the z table");
  AXS_CHECK_ARG(c0.history == 0 || bufs->drag, "history > 0 needs the drag buffer");
  axs_rigid_stepper* st = new (std::nothrow) axs_rigid_stepper();
  if (!st) {
    set_error("out of host memory");
    return AXS_ERR_ALLOC;
  }
  if (uneven && c0.precision == AXS_F64)
    st->impl = new (std::nothrow) RaggedStepper<double>();
  else if (uneven)
    st->impl = new (std::nothrow) RaggedStepper<float>();
  else if (c0.precision == AXS_F64)
    st->impl = new (std::nothrow) RigidStepper<double>();
  else
    st->impl = new (std::nothrow) RigidStepper<float>();
  if (!st->impl) {
    delete st;
    set_error("out of host memory");
    return AXS_ERR_ALLOC;
  }
  std::vector<RigidConfig> cs((size_t)batch);
  for (int q = 0; q < batch; ++q) cs[q] = to_internal(cfgs[q]);
  RigidBuffers b{bufs->vorticity, bufs->psi,    bufs->u_z,     bufs->u_r,    bufs->u_pen_z,
                 bufs->u_pen_r,   bufs->char_func, bufs->work0, bufs->work1,  bufs->avg_psi,
                 bufs->scalars,   bufs->drag,   bufs->r,       bufs->z,      bufs->kill_wz,
                 bufs->kill_wr};
  hipError_t e;
  if (!uneven)
    e = st->impl->init(cs.data(), batch, b);
  else if (c0.precision == AXS_F64)
    e = static_cast<RaggedStepper<double>*>(st->impl)->init_ragged(cs.data(), batch, b);
  else
    e = static_cast<RaggedStepper<float>*>(st->impl)->init_ragged(cs.data(), batch, b);
  if (e != hipSuccess) {
    set_error(std::string("rigid stepper: ") + hipGetErrorString(e));
    delete st;
    return AXS_ERR_HIP;
  }
  *out = st;
  return AXS_OK;
}

axs_status axs_rigid_stepper_create(const axs_rigid_config* cfg, const axs_rigid_buffers* bufs,
                                    axs_rigid_stepper** out) {
  return axs_rigid_batch_create(cfg, 1, bufs, out);
}

axs_status axs_rigid_stepper_destroy(axs_rigid_stepper* st) {
  delete st;
  return AXS_OK;
}

axs_status axs_rigid_stepper_run(axs_rigid_stepper* st, int nsteps, int use_graph, void* stream) {
  AXS_CHECK_ARG(st, "null stepper");
  AXS_CHECK_ARG(nsteps >= 0, "negative step count");
  AXS_CHECK_ARG(!use_graph || stream, "graph capture needs a non-null stream");
  AXS_HIP(st->impl->run(nsteps, use_graph != 0, as_stream(stream)));
  return AXS_OK;
}

axs_status axs_rigid_stepper_graph_launches(const axs_rigid_stepper* st, int64_t* chunks,
                                            int64_t* singles) {
  AXS_CHECK_ARG(st && chunks && singles, "null argument");
  *chunks = st->impl->chunk_launches;
  *singles = st->impl->step_launches;
  return AXS_OK;
}

int axs_rigid_stepper_stage_count(void) { return kRigidStages; }

const char* axs_rigid_stepper_stage_name(int stage) {
  return (stage >= 0 && stage < kRigidStages) ? kRigidStageNames[stage] : "";
}

axs_status axs_rigid_stepper_run_stage(axs_rigid_stepper* st, int stage, int reps, void* stream) {
  AXS_CHECK_ARG(st, "null stepper");
  AXS_CHECK_ARG(stage >= 0 && stage < kRigidStages, "bad stage");
  AXS_CHECK_ARG(reps >= 0, "negative repetition count");
  for (int q = 0; q < reps; ++q) AXS_HIP(st->impl->enqueue_stage(stage, as_stream(stream)));
  return AXS_OK;
}

axs_status axs_rigid_stepper_velocity(axs_rigid_stepper* st, void* stream) {
  AXS_CHECK_ARG(st, "null stepper");
  AXS_HIP(st->impl->velocity_out(as_stream(stream)));
  return AXS_OK;
}

axs_status axs_rigid_stepper_reset_scalars(axs_rigid_stepper* st, void* stream) {
  AXS_CHECK_ARG(st, "null stepper");
  init_scalars_kernel<<<st->impl->batch, 64, 0, as_stream(stream)>>>(
      st->impl->buf.scalars, st->impl->slots, st->impl->dcfg);
  AXS_LAUNCH_CHECK("init_scalars");
  return AXS_OK;
}

}  // extern "C"
