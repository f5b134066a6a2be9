// Fused stencil passes of the device-resident step (row A16).
//
//   velocity_max   u from psi (+ free stream) on 16 x 64 tiles ->
//                  max(|u_z| + |u_r|), folded with one atomicMax per block
//                  into one of kMaxSlots device slots (order-free, so the
//                  result is deterministic; slots on separate lines keep
//                  the atomics from serialising on one address).
//   step_fused     everything from the penalisation to the end of the step
//                  on a 16 x 54 output tile with a 6-cell halo:
//                    dt rule from the slot maximum (flow_past_sphere.py:
//                    150-153, spheroid_streaming.py:139-201),
//                    u from psi (compute_velocity_from_psi.py:4-17),
//                    Brinkman (brinkmann_penalize.py:4-16),
//                    kill taper of the stored field (kill_boundary_
//                    vorticity_sine.py:4-27), omega += curl(u_pen - u)
//                    (compute_vorticity_from_velocity.py:4-13),
//                    force partial sums (compute_forces.py:14), psi average,
//                    ENO3 step (advection_timestep.py:45-54, in-index axis
//                    mirror) and both RK2 stages (diffusion_RK2.py:4-44);
//                  reads psi, chi and the stored omega, writes only the new
//                  omega (into the other buffer of a ping-pong pair).
//   velocity_out   u_pen of the last step, on demand (not part of a step).
//   finalize       completes a step (fixed-order force sum, t += dt, it += 1)
//                  as an extra workgroup of the next step's first pass, or
//                  standalone at the end of a run.
// No per-block device-scope fences: the L2s of the XCDs are not coherent, so
// such a fence costs an L2 write-back; kernel boundaries order the passes.
//
// Per-cell arithmetic and its order are those of the single-purpose kernels.
// Divisions by dx^2, 2 dx and r_i use div_by (correctly rounded in FP64, so
// the results are unchanged; the reciprocal product in FP32); ENO3 faces shared by two cells are evaluated once
// (the backward face of a cell is the forward face of its lower neighbour,
// bit for bit: same comparison, same products, same sums).
#include "axs_fused.hpp"

#include <algorithm>

namespace axs {

// ------------------------------------------------------------ helpers

__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off);
  return v;
}

// u_z (with the free stream) and u_r at global (i, j); p points at that
// cell in a row-major array of row stride ld (global psi or an LDS tile)
// (compute_velocity_from_psi.py:10-17, one-sided at the edges).
template <typename T>
__device__ __forceinline__ void velocity_rn(const T* p, int ld, int i, int j, int nr, int nz,
                                            T two_dx, T inv2dx, T ri, T inv_ri, T ufree, T& uz,
                                            T& ur) {
  T nzv, nrv;
  if (i == 0)
    nzv = -p[2 * ld] + T(4) * p[ld] - T(3) * p[0];
  else if (i == nr - 1)
    nzv = p[-2 * ld] - T(4) * p[-ld] + T(3) * p[0];
  else
    nzv = p[ld] - p[-ld];
  if (j == 0)
    nrv = -(-p[2] + T(4) * p[1] - T(3) * p[0]);
  else if (j == nz - 1)
    nrv = -(p[-2] - T(4) * p[-1] + T(3) * p[0]);
  else
    nrv = -(p[1] - p[-1]);
  uz = div_by(div_by(nzv, two_dx, inv2dx), ri, inv_ri) + ufree;
  ur = div_by(div_by(nrv, two_dx, inv2dx), ri, inv_ri);
}

// (u + ldc U) / (1 + ldc); outside the body (ldc == 0) the quotient is the
// numerator itself, so the division is skipped without changing a bit.
template <typename T>
__device__ __forceinline__ T brinkman(T u, T ldc, T Uv) {
  T t = u + ldc * Uv;
  if (ldc != T(0)) t = t / (T(1) + ldc);
  return t;
}

__device__ __forceinline__ double load_slot(const double* p) {
  const unsigned long long b = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p),
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double(static_cast<long long>(b));
}

// Kill taper (kill_boundary_vorticity_sine.py:4-27): z ends of every row
// from columns w-1 and nz-w, then rows nr-w..nr-1 scaled copies of row
// nr-w, then row 0 = 0.  The tapered value of (i, j) is a weight times one
// stored value, the one at taper_source(i, j).
__device__ __forceinline__ int64_t taper_source(int i, int j, int nr, int nz, int w) {
  const int srow = (i >= nr - w) ? nr - w : i;
  const int jsrc = (j < w) ? w - 1 : ((j >= nz - w) ? nz - w : j);
  return (int64_t)srow * nz + jsrc;
}
template <typename T>
__device__ __forceinline__ T taper_apply(T v, int i, int j, int nr, int nz, int w,
                                         const double* __restrict__ wz,
                                         const double* __restrict__ wr) {
  T t = v;
  if (j < w)
    t = T(wz[j]) * v;
  else if (j >= nz - w)
    t = T(wz[w + (j - (nz - w))]) * v;
  if (i >= nr - w) t = T(wr[i - (nr - w)]) * t;
  if (i == 0) t = T(0);
  return t;
}

template <typename T>
__device__ __forceinline__ T eno3_face_f(T fm1, T f0, T fp1, T fp2, T um1, T u0, T up1, T up2) {
  // conservative face between cells 0 and +1 (advection_flux.py:32-44)
  return (u0 > -up1) ? (T(1) / T(3)) * fp1 * up1 + (T(5) / T(6)) * f0 * u0 - (T(1) / T(6)) * fm1 * um1
                     : (T(1) / T(3)) * f0 * u0 + (T(5) / T(6)) * fp1 * up1 - (T(1) / T(6)) * fp2 * up2;
}
template <typename T>
__device__ __forceinline__ T eno3_face_b(T fm2, T fm1, T f0, T fp1, T um2, T um1, T u0, T up1) {
  // conservative face between cells -1 and 0 (advection_flux.py:61-73)
  return (u0 > -um1) ? (T(1) / T(3)) * f0 * u0 + (T(5) / T(6)) * fm1 * um1 - (T(1) / T(6)) * fm2 * um2
                     : (T(1) / T(3)) * fm1 * um1 + (T(5) / T(6)) * f0 * u0 - (T(1) / T(6)) * fp1 * up1;
}

// ------------------------------------------------------------ velocity_max

constexpr int kVTZ = 64, kVTR = 16;  // tile of the one-grid kernel (TR = kVTR)
constexpr int kVPZ = kVTZ + 4;       // psi tile columns (halo 2)

// Tile (bx, by), TR rows by kVTZ columns (TR = 8, 16 or 32), of one case
// (psi, sc, slots at the case); the maximum goes to slot `slot`.
template <typename T, int TR, bool STREAM = false>
__device__ __forceinline__ void velocity_max_tile(const T* __restrict__ psi,
                                                  const double* __restrict__ r, int nr, int nz,
                                                  T dx, const double* __restrict__ sc,
                                                  double* __restrict__ slots,
                                                  const RigidConfig& cfg, int bx, int by,
                                                  int slot) {
  static_assert(TR == 8 || TR == 16 || TR == 32, "tile rows");
  constexpr int kVPR = TR + 4;  // psi tile rows (halo 2)
  __shared__ T sp[kVPR][kVPZ];
  __shared__ T sri[TR], sinv[TR];
  __shared__ double wm[4];
  const int tid = threadIdx.x;
  const int j0 = bx * kVTZ, i0 = by * TR;
  const T two_dx = T(2) * dx, inv2dx = T(1) / two_dx;
  const T ufree = T(cfg.U0 * sc[kScPrefac]);
  double m = 0.0;
  const bool interior = i0 >= 1 && i0 + TR <= nr - 1 && j0 >= 1 && j0 + kVTZ <= nz - 1;
  if (STREAM && interior) {
    // no cell of the tile lies on a domain edge: each thread walks TR / 4
    // rows of one column, reading psi and its four neighbours directly (the
    // neighbour rows and columns come from L1/L2), no LDS staging; central
    // differences only, the same arithmetic as the staged path below
    constexpr int RS = TR / 4;
    const int b = tid & 63, g = tid >> 6;
    const int ia = i0 + g * RS;
    const T* col = psi + (int64_t)ia * nz + (j0 + b);
    T c[RS + 2], lf[RS], rt[RS];
#pragma unroll
    for (int k = 0; k < RS + 2; ++k) c[k] = col[(int64_t)(k - 1) * nz];
#pragma unroll
    for (int k = 0; k < RS; ++k) {
      lf[k] = col[(int64_t)k * nz - 1];
      rt[k] = col[(int64_t)k * nz + 1];
    }
#pragma unroll
    for (int k = 0; k < RS; ++k) {
      const T ri = T(r[ia + k]);
      const T rinv = T(1) / ri;
      const T uz = div_by(div_by(c[k + 2] - c[k], two_dx, inv2dx), ri, rinv) + ufree;
      const T ur = div_by(div_by(-(rt[k] - lf[k]), two_dx, inv2dx), ri, rinv);
      m = fmax(m, double(fabs(uz) + fabs(ur)));
    }
  } else {
  constexpr int NP = (kVPR * kVPZ + 255) / 256;
  T v[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const int e = tid + q * 256;
    const int a = e / kVPZ, b = e % kVPZ;
    const int i = i0 - 2 + a, j = j0 - 2 + b;
    v[q] = T(0);
    if (e < kVPR * kVPZ && i >= 0 && i < nr && j >= 0 && j < nz) v[q] = psi[(int64_t)i * nz + j];
  }
  if (tid < TR) {
    const int i = i0 + tid;
    const T ri = (i < nr) ? T(r[i]) : T(1);
    sri[tid] = ri;
    sinv[tid] = T(1) / ri;
  }
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const int e = tid + q * 256;
    if (e < kVPR * kVPZ) sp[e / kVPZ][e % kVPZ] = v[q];
  }
  __syncthreads();
  if (interior) {
    // no cell of the tile lies on a domain edge: central differences only
    // (the formulas velocity_rn takes away from the edges)
#pragma unroll
    for (int q = 0; q < (TR * kVTZ) / 256; ++q) {
      const int e = tid + q * 256;
      const int a = e / kVTZ, b = e % kVTZ;
      const T* p = &sp[a + 2][b + 2];
      const T uz = div_by(div_by(p[kVPZ] - p[-kVPZ], two_dx, inv2dx), sri[a], sinv[a]) + ufree;
      const T ur = div_by(div_by(-(p[1] - p[-1]), two_dx, inv2dx), sri[a], sinv[a]);
      m = fmax(m, double(fabs(uz) + fabs(ur)));
    }
  } else {
#pragma unroll
    for (int q = 0; q < (TR * kVTZ) / 256; ++q) {
      const int e = tid + q * 256;
      const int a = e / kVTZ, b = e % kVTZ;
      const int i = i0 + a, j = j0 + b;
      if (i < nr && j < nz) {
        T uz, ur;
        velocity_rn<T>(&sp[a + 2][b + 2], kVPZ, i, j, nr, nz, two_dx, inv2dx, sri[a], sinv[a],
                       ufree, uz, ur);
        m = fmax(m, double(fabs(uz) + fabs(ur)));
      }
    }
  }
  }  // staged path
  m = wave_max(m);
  if ((tid & 63) == 0) wm[tid >> 6] = m;
  __syncthreads();
  if (tid == 0) {
    m = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
    atomicMax(reinterpret_cast<unsigned long long*>(slots + slot * kSlotStride),
              static_cast<unsigned long long>(__double_as_longlong(m)));
  }
}

template <typename T>
__global__ void __launch_bounds__(256) velocity_max_kernel(const T* __restrict__ psi,
                                                           const double* __restrict__ r, int nr,
                                                           int nz, T dx,
                                                           const double* __restrict__ sc,
                                                           double* __restrict__ slots,
                                                           const RigidConfig* __restrict__ cfgs) {
  const int cb = blockIdx.z;  // case of a batch
  velocity_max_tile<T, kVTR, true>(psi + (int64_t)cb * nr * nz, r, nr, nz, dx,
                       sc + (int64_t)cb * AXS_RIGID_SCALARS, slots + (int64_t)cb * kSlotsPerCase,
                       cfgs[cb], blockIdx.x, blockIdx.y,
                       (blockIdx.y * gridDim.x + blockIdx.x) % kMaxSlots);
}

template <typename T, int TR>
__global__ void __launch_bounds__(256) velocity_max_ragged_kernel(
    const T* __restrict__ psi, const double* __restrict__ sc, double* __restrict__ slots,
    const RigidConfig* __restrict__ cfgs, const RaggedGroup* __restrict__ grp, int ngrp) {
  const int b = blockIdx.x;
  const RaggedGroup& G = grp[ragged_group<&RaggedGroup::vt0>(grp, ngrp, b)];
  const int q = b - G.vt0, cl = q / G.vtiles, t = q % G.vtiles;
  const int cb = G.case0 + cl;
  velocity_max_tile<T, TR>(psi + G.cell0 + (int64_t)cl * G.nr * G.nz, G.r, G.nr, G.nz, T(G.dx),
                       sc + (int64_t)cb * AXS_RIGID_SCALARS, slots + (int64_t)cb * kSlotsPerCase,
                       cfgs[cb], t % G.vz, t / G.vz, t % kMaxSlots);
}

// ------------------------------------------------------------ step_fused

// Output tile kFR x kFZ.  Regions by halo around it: R6 psi, R5 velocity and
// penalisation, R4 ENO3 input (f, u_pen), R2 ENO3 output, R1 RK2 stage 1.
constexpr int kFZ = 54, kFR = 16;
constexpr int kR6Z = kFZ + 12, kR6R = kFR + 12;
constexpr int kR5Z = kFZ + 10, kR5R = kFR + 10;
constexpr int kR4Z = kFZ + 8, kR4R = kFR + 8;
constexpr int kR2Z = kFZ + 4, kR2R = kFR + 4;
constexpr int kR1Z = kFZ + 2, kR1R = kFR + 2;
// Thread strips: penalisation 4 x 64 (kVS owned R4 rows + one R5 row above
// and below), ENO3 rows kEH, stage 1 rows kSH, stage 2 rows kOH.
constexpr int kVS = 6, kEH = 5, kSH = 6, kOH = 4;
static_assert(kR5Z == 64 && 4 * kVS == kR4R, "penalisation strips cover R4 rows x R5 columns");
static_assert(kR2R % kEH == 0 && kR1R % kSH == 0 && kFR % kOH == 0, "strip heights");
static_assert((kR2R / kEH) * kR2Z <= 256 && (kR1R / kSH) * kR1Z <= 256 && (kFR / kOH) * kFZ <= 256,
              "one strip per thread");
// LDS buffers and what they hold phase by phase.
constexpr int kBufA = kR6R * kR6Z;  // psi (R6), then f (R4)
constexpr int kBufC = kR4R * kR4Z;  // u_pen_z (R4), then stage-1 out (R1)
constexpr int kBufD = kR4R * kR4Z;  // u_pen_r (R4), then ENO3 out (R2)
static_assert(kR4R * kR4Z <= kBufA && kR2R * kR2Z <= kBufD && kR1R * kR1Z <= kBufC, "aliases");

int fused_blocks(int nr, int nz) { return (int)(ceil_div(nz, kFZ) * ceil_div(nr, kFR)); }

// LDS of one step_fused block (passed to the tile body by pointer so both
// tile variants share one allocation).
template <typename T>
struct FusedLds {
  T* A;  // psi (R6), then f (R4)
  T* C;  // u_pen_z (R4), then stage-1 out (R1)
  T* D;  // u_pen_r (R4), then ENO3 out (R2)
  T *sri, *sinv, *sinv2;  // per R6 row: r_i, 1 / r_i, 1 / r_i^2
  double* srd;            // per R6 row: r_i (FP64, for the force)
  double* wpart;
  double* smax;  // slot maximum, then the step's dt, timer and reset flag
};

// One output tile.  IN: the tile and its 6-cell halo lie at least kill_width
// cells inside every domain edge (checked by the caller), so no cell needs a
// bounds test, a one-sided velocity formula, the kill taper, the axis mirror
// or a border exemption; the other tiles take every test.  RK2 false: the
// implicit-diffusion form of the drivers (flow_past_sphere.py:178-181), whose
// tile output is the ENO3 result (the diffusion solve follows as its own
// passes).
template <typename T, bool IN, bool RK2 = true>
__device__ __forceinline__ void step_fused_tile(
    const T* __restrict__ psi, const T* __restrict__ chi, const T* __restrict__ wsrc,
    T* __restrict__ wdst, T* __restrict__ avg, const double* __restrict__ r,
    const double* __restrict__ wz, const double* __restrict__ wr, int kw, int nr, int nz, T dx,
    double* __restrict__ sc, const double* __restrict__ slots, double* __restrict__ partials,
    const RigidConfig& cfg, const FusedLds<T>& L, int i0, int j0, int tile) {
  T(*sp)[kR6Z] = reinterpret_cast<T(*)[kR6Z]>(L.A);
  T(*sf)[kR4Z] = reinterpret_cast<T(*)[kR4Z]>(L.A);
  T(*sa)[kR2Z] = reinterpret_cast<T(*)[kR2Z]>(L.D);
  T(*sx)[kR4Z] = reinterpret_cast<T(*)[kR4Z]>(L.C);
  T(*st)[kR1Z] = reinterpret_cast<T(*)[kR1Z]>(L.C);
  T(*sy)[kR4Z] = reinterpret_cast<T(*)[kR4Z]>(L.D);
  const T* sri = L.sri;
  const T* sinv = L.sinv;
  const T* sinv2 = L.sinv2;
  const int tid = threadIdx.x;
  // penalisation strip of this thread: R5 column pb, R5 rows 6ps .. 6ps+7
  const int pb = tid & 63, ps = tid >> 6;
  const int pj = j0 - 5 + pb;
  const bool pjin = IN || (pj >= 0 && pj < nz);

  // ---- 1. loads: psi tile, chi and stored omega of the strip, slots, rows
  constexpr int NP = (kBufA + 255) / 256;
  T vp[NP];
  const T* ptile = psi + ((i0 - 6) * nz + (j0 - 6));  // interior tiles only
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const int e = tid + q * 256;
    const int a = e / kR6Z, b = e % kR6Z;
    vp[q] = T(0);
    if constexpr (IN) {
      if (e < kBufA) vp[q] = ptile[a * nz + b];
    } else {
      const int i = i0 - 6 + a, j = j0 - 6 + b;
      if (e < kBufA && i >= 0 && i < nr && j >= 0 && j < nz) vp[q] = psi[(int64_t)i * nz + j];
    }
  }
  T vc[kVS + 2], vw[kVS];
#pragma unroll
  for (int k = 0; k < kVS + 2; ++k) {
    const int i = i0 - 5 + kVS * ps + k;
    if constexpr (IN) {
      vc[k] = chi[i * nz + pj];
    } else {
      vc[k] = T(0);
      if (pjin && i >= 0 && i < nr) vc[k] = chi[(int64_t)i * nz + pj];
    }
  }
#pragma unroll
  for (int k = 0; k < kVS; ++k) {
    const int i = i0 - 4 + kVS * ps + k;
    if constexpr (IN) {
      vw[k] = wsrc[i * nz + pj];
    } else {
      vw[k] = T(0);
      if (pjin && i >= 0 && i < nr) vw[k] = wsrc[taper_source(i, pj, nr, nz, kw)];
    }
  }
  if (tid < 64) {
    const double m = wave_max(load_slot(slots + tid * kSlotStride));
    if (tid == 0) {
      // the dt rule once per block (not per thread: its temporaries would
      // overlap the tile loads still in flight)
      const DtScalars d = compute_dt(sc, m, cfg);
      L.smax[0] = m;
      L.smax[1] = d.dt;
      L.smax[2] = d.fr;
      L.smax[3] = d.reset;
    }
  }
  if (tid >= 64 && tid < 64 + kR6R) {
    const int i = i0 - 6 + (tid - 64);
    const double rd = (IN || (i >= 0 && i < nr)) ? r[i] : 1.0;
    const T ri = T(rd);
    L.srd[tid - 64] = rd;
    L.sri[tid - 64] = ri;
    L.sinv[tid - 64] = T(1) / ri;
    L.sinv2[tid - 64] = T(1) / (ri * ri);
  }
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const int e = tid + q * 256;
    if (e < kBufA) L.A[e] = vp[q];
  }
  // Does the body reach this tile's R5 region?  Where chi = 0, Brinkman
  // leaves u unchanged bit for bit (u + 0 U, no division), so u_pen - u = 0,
  // the penalisation curl is 0 and the force term is a signed zero.
  bool touches = false;
#pragma unroll
  for (int k = 0; k < kVS + 2; ++k) touches |= (vc[k] != T(0));
  const bool body = __syncthreads_or(touches) != 0;

  const DtScalars ds{L.smax[1], L.smax[2], L.smax[3]};
  const T two_dx = T(2) * dx, inv2dx = T(1) / two_dx;
  const T ufree = T(cfg.U0 * sc[kScPrefac]);
  const T lam_dt = T(cfg.brink_lam * ds.dt);
  const double Ud = sc[kScUpen];
  const T U = T(Ud);
  const T tdt = T(ds.dt);
  double acc = 0.0;
  if (IN && !body) {
    // ---- 2'. body-free interior tile: u on the R4 cells only (the ENO3
    //          velocity), f = omega + 0 (the zero curl, so signed zeros
    //          match the general path)
    if (pb >= 1 && pb <= kR4Z) {
#pragma unroll
      for (int k = 1; k <= kVS; ++k) {
        const int q5 = kVS * ps + k, a4 = q5 - 1;
        const T* p = &sp[q5 + 1][pb + 1];
        const T uz =
            div_by(div_by(p[kR6Z] - p[-kR6Z], two_dx, inv2dx), sri[q5 + 1], sinv[q5 + 1]) + ufree;
        const T ur = div_by(div_by(-(p[1] - p[-1]), two_dx, inv2dx), sri[q5 + 1], sinv[q5 + 1]);
        const T ldc = lam_dt * vc[k];
        const T pzv = brinkman(uz, ldc, U);
        const T prv = brinkman(ur, ldc, T(0));
        if (avg && q5 >= 5 && q5 < 5 + kFR && pb >= 5 && pb < 5 + kFZ) {
          const int c = (i0 - 5 + q5) * nz + pj;
          const T base = (ds.reset != 0.0) ? T(0) : avg[c];
          avg[c] = base + p[0] * tdt;
        }
        sx[a4][pb - 1] = pzv;
        sy[a4][pb - 1] = prv;
        vw[k - 1] = vw[k - 1] + T(0);
      }
    }
    __syncthreads();  // the psi tile (bufA) is read above and overwritten below
    if (pb >= 1 && pb <= kR4Z) {
#pragma unroll
      for (int k = 1; k <= kVS; ++k) sf[kVS * ps + k - 1][pb - 1] = vw[k - 1];
    }
    __syncthreads();
  } else {
  T dz[kVS + 2];  // u_pen_z - u_z on the strip's R5 rows
  T dr[kVS];      // u_pen_r - u_r on the strip's R4 rows
#pragma unroll
  for (int k = 0; k < kVS + 2; ++k) {
    const int q5 = kVS * ps + k;  // R5 row
    const int i = i0 - 5 + q5;
    T dzv = T(0), drv = T(0), pzv = T(0), prv = T(0);
    if (IN || (pjin && i >= 0 && i < nr)) {
      const T* p = &sp[q5 + 1][pb + 1];
      T uz, ur;
      if constexpr (IN) {
        // compute_velocity_from_psi.py:10-17, central differences
        uz = div_by(div_by(p[kR6Z] - p[-kR6Z], two_dx, inv2dx), sri[q5 + 1], sinv[q5 + 1]) + ufree;
        ur = div_by(div_by(-(p[1] - p[-1]), two_dx, inv2dx), sri[q5 + 1], sinv[q5 + 1]);
      } else {
        velocity_rn<T>(p, kR6Z, i, pj, nr, nz, two_dx, inv2dx, sri[q5 + 1], sinv[q5 + 1], ufree,
                       uz, ur);
      }
      const T ldc = lam_dt * vc[k];
      pzv = brinkman(uz, ldc, U);
      prv = brinkman(ur, ldc, T(0));
      dzv = pzv - uz;
      drv = prv - ur;
      // output tile cells: R5 rows 5..5+kFR-1, columns 5..5+kFZ-1
      if (k >= 1 && k <= kVS && q5 >= 5 && q5 < 5 + kFR && pb >= 5 && pb < 5 + kFZ &&
          (IN || (i < nr && pj < nz))) {
        if (avg) {
          const int64_t c = (int64_t)i * nz + pj;
          const T base = (ds.reset != 0.0) ? T(0) : avg[c];
          avg[c] = base + p[0] * tdt;
        }
        acc += (L.srd[q5 + 1] * double(vc[k])) * (double(pzv) - Ud);
      }
    }
    dz[k] = dzv;
    if (k >= 1 && k <= kVS) {
      const int a4 = q5 - 1;
      dr[k - 1] = drv;
      if (pb >= 1 && pb <= kR4Z) {
        sx[a4][pb - 1] = pzv;
        sy[a4][pb - 1] = prv;
      }
    }
  }
  __syncthreads();

  // ---- 3. f = taper(omega) + curl(u_pen - u) on the strip's R4 cells; the
  //         z neighbours of a strip column are the neighbouring lanes
  T drp[kVS], drm[kVS];
#pragma unroll
  for (int k = 0; k < kVS; ++k) {
    drp[k] = __shfl_down(dr[k], 1);
    drm[k] = __shfl_up(dr[k], 1);
  }
  if (pb >= 1 && pb <= kR4Z) {
#pragma unroll
    for (int k = 1; k <= kVS; ++k) {
      const int a4 = kVS * ps + k - 1;
      const int i = i0 - 4 + a4;
      T fv = T(0);
      if constexpr (IN) {
        const T pv = div_by(drp[k - 1] - drm[k - 1], two_dx, inv2dx) -
                     div_by(dz[k + 1] - dz[k - 1], two_dx, inv2dx);
        fv = vw[k - 1] + pv;
      } else {
        if (pjin && i >= 0 && i < nr) {
          fv = taper_apply(vw[k - 1], i, pj, nr, nz, kw, wz, wr);
          if (i >= 1 && i <= nr - 2 && pj >= 1 && pj <= nz - 2) {
            const T pv = div_by(drp[k - 1] - drm[k - 1], two_dx, inv2dx) -
                         div_by(dz[k + 1] - dz[k - 1], two_dx, inv2dx);
            fv = fv + pv;
          }
        }
      }
      sf[a4][pb - 1] = fv;
    }
  }
  __syncthreads();
  }  // body path
  // the tile's force sum leaves the registers before the ENO3 and RK2
  // phases (the same wave reduction the end of the kernel would make)
  acc = wave_sum(acc);
  if ((tid & 63) == 0) L.wpart[tid >> 6] = acc;
  if constexpr (!IN) {
    if (i0 == 0) {
      // rows -4..-1 of the ENO3 input are the axis mirror of rows 3..0
      // (f odd, u_z even, u_r odd: advect_vorticity_via_eno3.py:28-33)
      for (int e = tid; e < 4 * kR4Z; e += 256) {
        const int m = e / kR4Z, c = e % kR4Z;
        sf[3 - m][c] = -sf[4 + m][c];
        sx[3 - m][c] = sx[4 + m][c];
        sy[3 - m][c] = -sy[4 + m][c];
      }
      __syncthreads();
    }
  }

  // ---- 4. ENO3 on R2 (advection_timestep.py:45-54); r faces slide through
  //         registers, each evaluated once
  const T s = -T(ds.dt / cfg.dx);
  const bool eno = tid < (kR2R / kEH) * kR2Z;
  const int eb = tid % kR2Z, ea0 = (tid / kR2Z) * kEH;
  T out[kEH];
  if (eno) {
    const int b = eb, a0 = ea0;
    const int ib = b + 2;  // R4 column
    // R4 rows a0+k .. a0+k+3 of this column slide through a window of four;
    // face m lies between R2 rows a0+m-1 and a0+m: cell k uses faces k and
    // k+1, so only two are live
    T w0 = sf[a0][ib], w1 = sf[a0 + 1][ib], w2 = sf[a0 + 2][ib], w3 = sf[a0 + 3][ib];
    T y0 = sy[a0][ib], y1 = sy[a0 + 1][ib], y2 = sy[a0 + 2][ib], y3 = sy[a0 + 3][ib];
    T face_lo = eno3_face_f(w0, w1, w2, w3, y0, y1, y2, y3);
    const int j = j0 - 2 + b;
    const bool jok = IN || (j >= 2 && j < nz - 2);
#pragma unroll
    for (int k = 0; k < kEH; ++k) {
      const int ia = a0 + k + 2;
      const int i = i0 - 2 + a0 + k;
      T v = w2;  // R4 row a0+k+2
      w0 = w1;
      w1 = w2;
      w2 = w3;
      w3 = sf[a0 + k + 4][ib];
      y0 = y1;
      y1 = y2;
      y2 = y3;
      y3 = sy[a0 + k + 4][ib];
      const T face_hi = eno3_face_f(w0, w1, w2, w3, y0, y1, y2, y3);
      if (IN || (jok && i >= 0 && i < nr - 2)) {
        const T fm2 = sf[ia][ib - 2], fm1 = sf[ia][ib - 1], fp1 = sf[ia][ib + 1],
                fp2 = sf[ia][ib + 2];
        const T xm2 = sx[ia][ib - 2], xm1 = sx[ia][ib - 1], x0 = sx[ia][ib],
                xp1 = sx[ia][ib + 1], xp2 = sx[ia][ib + 2];
        T flux = T(0) + s * eno3_face_f(fm1, v, fp1, fp2, xm1, x0, xp1, xp2);
        flux = flux - s * eno3_face_b(fm2, fm1, v, fp1, xm2, xm1, x0, xp1);
        flux = flux + s * face_hi;
        flux = flux - s * face_lo;
        v = v + flux;
      }
      out[k] = v;
      face_lo = face_hi;
    }
  }
  __syncthreads();  // u_pen_r (bufD) is read above and overwritten below
  if (eno) {
#pragma unroll
    for (int k = 0; k < kEH; ++k) sa[ea0 + k][eb] = out[k];
  }
  __syncthreads();

  if constexpr (!RK2) {
    // the ENO3 output of the tile's own cells is the step's result
    if (tid < (kFR / kOH) * kFZ) {
      const int b = tid % kFZ, a0 = (tid / kFZ) * kOH;
      const int j = j0 + b;
#pragma unroll
      for (int k = 0; k < kOH; ++k) {
        const int i = i0 + a0 + k;
        if (!IN && (j >= nz || i >= nr)) continue;
        wdst[(int64_t)i * nz + j] = sa[a0 + k + 2][b + 2];
      }
    }
  } else {
  // ---- 5. RK2 stage 1 on R1 (diffusion_RK2.py:9-26)
  const T dx2 = dx * dx, inv_dx2 = T(1) / dx2;
  const T half_nu_dt = T(0.5 * cfg.nu * ds.dt), nu_dt = T(cfg.nu * ds.dt);
  if (tid < (kR1R / kSH) * kR1Z) {
    const int b = tid % kR1Z, a0 = (tid / kR1Z) * kSH;
    const int eb = b + 1;  // R2 column
    T ws[kSH + 2];         // R2 rows a0 .. a0 + kSH + 1
#pragma unroll
    for (int k = 0; k < kSH + 2; ++k) ws[k] = sa[a0 + k][eb];
    const int j = j0 - 1 + b;
    const bool jok = IN || (j >= 1 && j <= nz - 2);
#pragma unroll
    for (int k = 0; k < kSH; ++k) {
      const int a = a0 + k, ea = a + 1, a6 = a + 5;
      const int i = i0 - 1 + a;
      T v = ws[k + 1];
      if (IN || (jok && i >= 1 && i <= nr - 2)) {
        const T up = ws[k + 2], dn = ws[k];
        const T Lp = div_by(up + dn + sa[ea][eb + 1] + sa[ea][eb - 1] - T(4) * v, dx2, inv_dx2) +
                     div_by(div_by(up - dn, two_dx, inv2dx), sri[a6], sinv[a6]) - v * sinv2[a6];
        v = v + half_nu_dt * Lp;
      }
      st[a][b] = v;  // bufC (u_pen_z) was last read before the previous barrier
    }
  }
  __syncthreads();

  // ---- 6. RK2 stage 2 on the output tile (diffusion_RK2.py:27-44)
  if (tid < (kFR / kOH) * kFZ) {
    const int b = tid % kFZ, a0 = (tid / kFZ) * kOH;
    const int tb = b + 1;  // R1 column
    T wt[kOH + 2];         // R1 rows a0 .. a0 + kOH + 1
#pragma unroll
    for (int k = 0; k < kOH + 2; ++k) wt[k] = st[a0 + k][tb];
    const int j = j0 + b;
    const bool jin = IN || j < nz, jok = IN || (j >= 1 && j <= nz - 2);
#pragma unroll
    for (int k = 0; k < kOH; ++k) {
      const int a = a0 + k, ta = a + 1, a6 = a + 6;
      const int i = i0 + a;
      if (!IN && (!jin || i >= nr)) continue;
      T v = sa[a + 2][b + 2];
      if (IN || (jok && i >= 1 && i <= nr - 2)) {
        const T tc = wt[k + 1], up = wt[k + 2], dn = wt[k];
        const T Lp = div_by(up + dn + st[ta][tb + 1] + st[ta][tb - 1] - T(4) * tc, dx2, inv_dx2) +
                     div_by(div_by(up - dn, two_dx, inv2dx), sri[a6], sinv[a6]) - tc * sinv2[a6];
        v = v + nu_dt * Lp;
      }
      if constexpr (IN)
        wdst[i * nz + j] = v;
      else
        wdst[(int64_t)i * nz + j] = v;
    }
  }
  }  // RK2

  // ---- force partial of this tile; block (0, 0) publishes the dt scalars
  //      (read back from LDS, not held in registers through the phases)
  __syncthreads();
  if (tid == 0) {
    partials[tile] = ((L.wpart[0] + L.wpart[1]) + L.wpart[2]) + L.wpart[3];
    if (tile == 0)
      store_dt_scalars(sc, DtScalars{L.smax[1], L.smax[2], L.smax[3]}, L.smax[0], cfg);
  }
}

// Waves per SIMD the fused kernels are compiled for.  FP64: at most 128
// VGPRs and under 40 KiB of LDS, four blocks per CU.  FP32 one-grid: the
// same bound (82 VGPRs, five blocks per CU; a one-grid launch of the
// precision form has two or three blocks per CU, so more occupancy buys
// nothing and spills cost 4%).  FP32 uneven batches: at most 72 VGPRs,
// seven blocks per CU (the spills this costs are cheaper than the lost
// occupancy; per batched step 367 us at five blocks, 371 at six, 335 at
// eight, 327 at seven).
template <typename T>
constexpr int kFusedWaves = 4;
template <typename T>
constexpr int kRaggedFusedWaves = sizeof(T) == 8 ? 4 : 7;

// 256 threads, waves and LDS as kFusedWaves says.
template <typename T, bool RK2>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kFusedWaves<T>))) step_fused_kernel(
    const T* __restrict__ psi, const T* __restrict__ chi, const T* __restrict__ wsrc,
    T* __restrict__ wdst, T* __restrict__ avg, const double* __restrict__ r,
    const double* __restrict__ wz, const double* __restrict__ wr, int kw, int nr, int nz, T dx,
    double* __restrict__ sc, const double* __restrict__ slots, double* __restrict__ partials,
    const RigidConfig* __restrict__ cfgs) {
  __shared__ T bufA[kBufA], bufC[kBufC], bufD[kBufD];
  __shared__ T sri[kR6R], sinv[kR6R], sinv2[kR6R];
  __shared__ double srd[kR6R];
  __shared__ double wpart[4];
  __shared__ double smax[4];
  const int cb = blockIdx.z;  // case of a batch
  const RigidConfig& cfg = cfgs[cb];
  const int64_t fo = (int64_t)cb * nr * nz;
  psi += fo;
  chi += fo;
  wsrc += fo;
  wdst += fo;
  if (avg) avg += fo;
  sc += (int64_t)cb * AXS_RIGID_SCALARS;
  slots += (int64_t)cb * kSlotsPerCase;
  partials += (int64_t)cb * gridDim.x * gridDim.y;
  const FusedLds<T> L{bufA, bufC, bufD, sri, sinv, sinv2, srd, wpart, smax};
  // XCD-aware tile order: workgroups are dealt to the 8 XCDs round-robin by
  // linear id; give each XCD a contiguous band of tiles so a tile's halo
  // rows are still in that XCD's L2 when its neighbours read them.
  const int gx = gridDim.x, ntile = gx * gridDim.y;
  int tile = blockIdx.y * gx + blockIdx.x;
  if ((ntile & 7) == 0) tile = (tile & 7) * (ntile >> 3) + (tile >> 3);
  const int j0 = (tile % gx) * kFZ, i0 = (tile / gx) * kFR;
  // interior: every R6 cell exists and every R4 cell is clear of the taper
  // (kill_width columns at each z end, kill_width rows below r_max, row 0)
  const int m = kw > 1 ? kw : 1;
  const bool interior = i0 - 6 >= m && i0 + kFR + 6 <= nr - m && j0 - 6 >= m &&
                        j0 + kFZ + 6 <= nz - m;
  if (interior)
    step_fused_tile<T, true, RK2>(psi, chi, wsrc, wdst, avg, r, wz, wr, kw, nr, nz, dx, sc, slots,
                             partials, cfg, L, i0, j0, tile);
  else
    step_fused_tile<T, false, RK2>(psi, chi, wsrc, wdst, avg, r, wz, wr, kw, nr, nz, dx, sc, slots,
                              partials, cfg, L, i0, j0, tile);
}

template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kRaggedFusedWaves<T>)))
step_fused_ragged_kernel(const T* __restrict__ psi, const T* __restrict__ chi,
                         const T* __restrict__ wsrc, T* __restrict__ wdst, T* __restrict__ avg,
                         int kw, double* __restrict__ sc, const double* __restrict__ slots,
                         double* __restrict__ partials, const RigidConfig* __restrict__ cfgs,
                         const RaggedGroup* __restrict__ grp, int ngrp) {
  __shared__ T bufA[kBufA], bufC[kBufC], bufD[kBufD];
  __shared__ T sri[kR6R], sinv[kR6R], sinv2[kR6R];
  __shared__ double srd[kR6R];
  __shared__ double wpart[4];
  __shared__ double smax[4];
  // the same XCD-aware order as the uniform launch, over all tiles of all cases
  const int nb = gridDim.x;
  int b = blockIdx.x;
  if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
  const RaggedGroup& G = grp[ragged_group<&RaggedGroup::ft0>(grp, ngrp, b)];
  const int q = b - G.ft0, cl = q / G.ftiles, tile = q % G.ftiles;
  const int cb = G.case0 + cl;
  const int nr = G.nr, nz = G.nz;
  const int64_t fo = G.cell0 + (int64_t)cl * nr * nz;
  const FusedLds<T> L{bufA, bufC, bufD, sri, sinv, sinv2, srd, wpart, smax};
  const int j0 = (tile % G.fz) * kFZ, i0 = (tile / G.fz) * kFR;
  const int m = kw > 1 ? kw : 1;
  const bool interior = i0 - 6 >= m && i0 + kFR + 6 <= nr - m && j0 - 6 >= m &&
                        j0 + kFZ + 6 <= nz - m;
  double* part = partials + G.part0 + (int64_t)cl * G.ftiles;
  if (interior)
    step_fused_tile<T, true>(psi + fo, chi + fo, wsrc + fo, wdst + fo, avg ? avg + fo : nullptr,
                             G.r, G.kill_wz, G.kill_wr, kw, nr, nz, T(G.dx),
                             sc + (int64_t)cb * AXS_RIGID_SCALARS,
                             slots + (int64_t)cb * kSlotsPerCase, part, cfgs[cb], L, i0, j0, tile);
  else
    step_fused_tile<T, false>(psi + fo, chi + fo, wsrc + fo, wdst + fo, avg ? avg + fo : nullptr,
                              G.r, G.kill_wz, G.kill_wr, kw, nr, nz, T(G.dx),
                              sc + (int64_t)cb * AXS_RIGID_SCALARS,
                              slots + (int64_t)cb * kSlotsPerCase, part, cfgs[cb], L, i0, j0, tile);
}

// ------------------------------------------------------------ velocity_out

// u_pen of the last step from psi and chi with that step's dt, ramp factor
// and penalisation velocity (the values step_fused used, bit for bit).
template <typename T>
__global__ void __launch_bounds__(256) velocity_out_kernel(const T* __restrict__ psi,
                                                           const T* __restrict__ chi,
                                                           T* __restrict__ pz, T* __restrict__ pr,
                                                           const double* __restrict__ r, int nr,
                                                           int nz, T dx,
                                                           const double* __restrict__ sc,
                                                           const RigidConfig* __restrict__ cfgs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  const int cb = blockIdx.z;  // case of a batch
  const RigidConfig& cfg = cfgs[cb];
  const int64_t fo = (int64_t)cb * nr * nz;
  psi += fo;
  chi += fo;
  pz += fo;
  pr += fo;
  sc += (int64_t)cb * AXS_RIGID_SCALARS;
  const T two_dx = T(2) * dx, inv2dx = T(1) / two_dx;
  const T ri = T(r[i]);
  const int64_t c = (int64_t)i * nz + j;
  T uz, ur;
  velocity_rn<T>(psi + c, nz, i, j, nr, nz, two_dx, inv2dx, ri, T(1) / ri,
                 T(cfg.U0 * sc[kScPrefacStep]), uz, ur);
  const T ldc = T(sc[kScLamDt]) * chi[c];
  pz[c] = brinkman(uz, ldc, T(sc[kScUpenStep]));
  pr[c] = brinkman(ur, ldc, T(0));
}

// ------------------------------------------------------------ finalize

__global__ void __launch_bounds__(1024) finalize_step_kernel(const double* __restrict__ partials,
                                                             int count, double* __restrict__ slots,
                                                             double* __restrict__ sc,
                                                             double* __restrict__ drag,
                                                             const RigidConfig* __restrict__ cfgs) {
  __shared__ double red[1024];
  const int cb = blockIdx.x;  // one block per case
  const RigidConfig& cfg = cfgs[cb];
  finalize_pending(partials + (int64_t)cb * count, count, slots + (int64_t)cb * kSlotsPerCase,
                   sc + (int64_t)cb * AXS_RIGID_SCALARS,
                   drag ? drag + (int64_t)cb * cfg.history : nullptr, cfg, red);
}

__global__ void __launch_bounds__(1024) finalize_ragged_kernel(
    const double* __restrict__ partials, double* __restrict__ slots, double* __restrict__ sc,
    double* __restrict__ drag, const RigidConfig* __restrict__ cfgs,
    const RaggedGroup* __restrict__ grp, int ngrp) {
  __shared__ double red[1024];
  const int cb = blockIdx.x;  // one block per case
  const RaggedGroup& G = grp[ragged_group<&RaggedGroup::case0>(grp, ngrp, cb)];
  const RigidConfig& cfg = cfgs[cb];
  finalize_pending(partials + G.part0 + (int64_t)(cb - G.case0) * G.ftiles, G.ftiles,
                   slots + (int64_t)cb * kSlotsPerCase, sc + (int64_t)cb * AXS_RIGID_SCALARS,
                   drag ? drag + (int64_t)cb * cfg.history : nullptr, cfg, red);
}

// ------------------------------------------------------------ launchers

template <typename T>
hipError_t launch_velocity_max(const T* psi, const double* r, int nr, int nz, T dx,
                               const double* sc, double* slots, const RigidConfig* cfgs,
                               int batch, hipStream_t s) {
  velocity_max_kernel<T><<<dim3(ceil_div(nz, kVTZ), ceil_div(nr, kVTR), batch), 256, 0, s>>>(
      psi, r, nr, nz, dx, sc, slots, cfgs);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_step_fused(const T* psi, const T* chi, const T* wsrc, T* wdst, T* avg,
                             const double* r, const double* wz, const double* wr, int kw, int nr,
                             int nz, T dx, double* sc, const double* slots, double* partials,
                             const RigidConfig* cfgs, int batch, hipStream_t s, bool rk2) {
  const dim3 grid(ceil_div(nz, kFZ), ceil_div(nr, kFR), batch);
  if (rk2)
    step_fused_kernel<T, true><<<grid, 256, 0, s>>>(psi, chi, wsrc, wdst, avg, r, wz, wr, kw, nr,
                                                    nz, dx, sc, slots, partials, cfgs);
  else
    step_fused_kernel<T, false><<<grid, 256, 0, s>>>(psi, chi, wsrc, wdst, avg, r, wz, wr, kw, nr,
                                                     nz, dx, sc, slots, partials, cfgs);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_velocity_out(const T* psi, const T* chi, T* pz, T* pr, const double* r, int nr,
                               int nz, T dx, const double* sc, const RigidConfig* cfgs, int batch,
                               hipStream_t s) {
  velocity_out_kernel<T><<<dim3(ceil_div(nz, 256), nr, batch), 256, 0, s>>>(psi, chi, pz, pr, r,
                                                                             nr, nz, dx, sc, cfgs);
  return hipGetLastError();
}

hipError_t launch_finalize_step(const double* partials, int count, double* slots, double* sc,
                                double* drag, const RigidConfig* cfgs, int batch, hipStream_t s) {
  finalize_step_kernel<<<batch, 1024, 0, s>>>(partials, count, slots, sc, drag, cfgs);
  return hipGetLastError();
}

int fused_tiles_z(int nz) { return (int)ceil_div(nz, kFZ); }
int vmax_tiles_z(int nz) { return (int)ceil_div(nz, kVTZ); }
int vmax_tiles(int nr, int nz, int tr) { return (int)(ceil_div(nz, kVTZ) * ceil_div(nr, tr)); }

template <typename T>
hipError_t launch_velocity_max_ragged(const T* psi, const double* sc, double* slots,
                                      const RigidConfig* cfgs, const RaggedGroup* grp, int ngrp,
                                      int tiles, int tr, hipStream_t s) {
  if (tr == 8)
    velocity_max_ragged_kernel<T, 8><<<tiles, 256, 0, s>>>(psi, sc, slots, cfgs, grp, ngrp);
  else if (tr == 16)
    velocity_max_ragged_kernel<T, 16><<<tiles, 256, 0, s>>>(psi, sc, slots, cfgs, grp, ngrp);
  else if (tr == 32)
    velocity_max_ragged_kernel<T, 32><<<tiles, 256, 0, s>>>(psi, sc, slots, cfgs, grp, ngrp);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <typename T>
hipError_t launch_step_fused_ragged(const T* psi, const T* chi, const T* wsrc, T* wdst, T* avg,
                                    int kw, double* sc, const double* slots, double* partials,
                                    const RigidConfig* cfgs, const RaggedGroup* grp, int ngrp,
                                    int tiles, hipStream_t s) {
  step_fused_ragged_kernel<T><<<tiles, 256, 0, s>>>(psi, chi, wsrc, wdst, avg, kw, sc, slots,
                                                    partials, cfgs, grp, ngrp);
  return hipGetLastError();
}

hipError_t launch_finalize_ragged(const double* partials, double* slots, double* sc, double* drag,
                                  const RigidConfig* cfgs, const RaggedGroup* grp, int ngrp,
                                  int ncases, hipStream_t s) {
  finalize_ragged_kernel<<<ncases, 1024, 0, s>>>(partials, slots, sc, drag, cfgs, grp, ngrp);
  return hipGetLastError();
}

#define AXS_FUSED_INST(T)                                                                     \
  template hipError_t launch_velocity_max<T>(const T*, const double*, int, int, T,          \
                                             const double*, double*, const RigidConfig*, int, \
                                             hipStream_t);                                   \
  template hipError_t launch_step_fused<T>(const T*, const T*, const T*, T*, T*,            \
                                           const double*, const double*, const double*, int, \
                                           int, int, T, double*, const double*, double*,    \
                                           const RigidConfig*, int, hipStream_t, bool);      \
  template hipError_t launch_velocity_out<T>(const T*, const T*, T*, T*, const double*, int, \
                                             int, T, const double*, const RigidConfig*, int, \
                                             hipStream_t);
AXS_FUSED_INST(double)
AXS_FUSED_INST(float)
template hipError_t launch_velocity_max_ragged<double>(const double*, const double*, double*,
                                                       const RigidConfig*, const RaggedGroup*,
                                                       int, int, int, hipStream_t);
template hipError_t launch_velocity_max_ragged<float>(const float*, const double*, double*,
                                                      const RigidConfig*, const RaggedGroup*, int,
                                                      int, int, hipStream_t);
template hipError_t launch_step_fused_ragged<double>(const double*, const double*, const double*,
                                                     double*, double*, int, double*,
                                                     const double*, double*, const RigidConfig*,
                                                     const RaggedGroup*, int, int, hipStream_t);
template hipError_t launch_step_fused_ragged<float>(const float*, const float*, const float*,
                                                    float*, float*, int, double*, const double*,
                                                    double*, const RigidConfig*,
                                                    const RaggedGroup*, int, int, hipStream_t);
#undef AXS_FUSED_INST

}  // namespace axs
