// Narrow-band signed-distance reinitialisation (row A15).
//
// The reference calls scikit-fmm 2022.8.15 (skfmm.distance with order 2,
// soft_sphere_in_taylor_green_vortex.py:159-162), a sequential fast
// marching method.  Marching in order of distance does not map onto a GPU;
// here the same discrete equations are solved by Jacobi iteration, which is
// data parallel:
//   init    interface cells (phi == 0, or a sign change towards an axis
//           neighbour) are frozen at the distance interpolated linearly to
//           the crossing along each axis, d_a = dx phi / (phi - phi_nb),
//           combined as 1 / sqrt(sum_a 1 / d_a^2), with the sign of phi;
//   sweep   every other cell takes the second-order upwind eikonal update
//           |grad d| = 1 from its known axis neighbours (values relative to
//           its own side of the interface), replacing the iterate (Jacobi,
//           ping-pong buffers); at the fixed point every cell satisfies its
//           upwind update from its converged neighbours;
//   final   |d| > narrow is masked and keeps phi.
// FP64 with the operation order of the CPU restatement used by the tests.
#include "axs_reinit.hpp"
#include "axs_softbatch.hpp"

#include <algorithm>
#include <cmath>

namespace axs {

constexpr double kBig = 1e300;

int reinit_iterations(double narrow, double dx) {
  return 2 * static_cast<int>(std::ceil(narrow / dx)) + 8;
}

// The sweeps run on 8 x 32 tiles (kReinitTileRows x kReinitTileCols).
__host__ __device__ __forceinline__ int reinit_tile_cols(int nz) {
  return (nz + kReinitTileCols - 1) / kReinitTileCols;
}
__host__ __device__ __forceinline__ int reinit_tile_count(int nr, int nz) {
  return ((nr + kReinitTileRows - 1) / kReinitTileRows) * reinit_tile_cols(nz);
}

size_t reinit_tile_flag_bytes(int nr, int nz) { return 4 * (size_t)reinit_tile_count(nr, nz); }

// Tile flags (ReinitWork::tiles, four rows of one byte per tile): row p
// (p = launch parity) = some cell of the tile changed during launch k,
// row 2 + p = some cell changed in launch k's last sweep.  Before the first
// launch they read as launch -1 (row 1: nothing, row 3: the tile holds an
// interface cell) and launch -2 (row 2: nothing).
//
// Both iterate buffers start as the initial iterate (kBig off the
// interface), with the interface cells marked changed in the first and
// nothing in the second: a cell none of whose neighbours is an interface
// cell has only kBig neighbours, which leave it kBig, so marking the
// interface cells alone starts the sweeps as marking every cell would.
__device__ __forceinline__ void reinit_init_at(const double* __restrict__ phi,
                                               double* __restrict__ D, double* __restrict__ D1,
                                               double* __restrict__ dist,
                                               uint8_t* __restrict__ frozen,
                                               uint8_t* __restrict__ ch0, uint8_t* __restrict__ ch1,
                                               uint8_t* __restrict__ tiles, int nr, int nz,
                                               double dx, int i, int j) {
  const int64_t c = (int64_t)i * nz + j;
  const double p = phi[c];
  double inv = 0.0;
  bool border = false;
#pragma unroll
  for (int axis = 0; axis < 2; ++axis) {
    double dmin = HUGE_VAL;
#pragma unroll
    for (int s = -1; s <= 1; s += 2) {
      const int ii = axis == 0 ? i + s : i, jj = axis == 0 ? j : j + s;
      if (ii < 0 || ii >= nr || jj < 0 || jj >= nz) continue;
      const double nb = phi[(int64_t)ii * nz + jj];
      if (p * nb < 0.0) dmin = fmin(dmin, dx * p / (p - nb));
    }
    if (dmin < HUGE_VAL) {
      border = true;
      inv = inv + 1.0 / dmin / dmin;
    }
  }
  double dv = 0.0;
  bool fr = false;
  if (border) {
    const double mag = sqrt(1.0 / inv);
    dv = (p < 0.0) ? -mag : mag;
    fr = true;
  }
  if (p == 0.0) {
    dv = 0.0;
    fr = true;
  }
  dist[c] = dv;
  frozen[c] = fr ? 1 : 0;
  D[c] = fr ? dv : kBig;
  D1[c] = fr ? dv : kBig;
  ch0[c] = fr ? 1 : 0;
  ch1[c] = 0;
  if (fr)
    tiles[3 * (size_t)reinit_tile_count(nr, nz) + (i / kReinitTileRows) * reinit_tile_cols(nz) +
          j / kReinitTileCols] = 1;
}

__global__ void reinit_init_kernel(const double* __restrict__ phi, ReinitWork w, int nr, int nz,
                                   double dx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  reinit_init_at(phi, w.d0, w.d1, w.dist, w.frozen, w.ch0, w.ch1, w.tiles, nr, nz, dx, i, j);
}

// Upwind eikonal update of cell (i, j) from the iterate D (kBig = unknown),
// neighbour values taken relative to the cell's side.  Per axis: v1 = the
// smaller known neighbour, v2 = the next cell beyond it when known and not
// larger; second-order term (9/4)(x - (4 v1 - v2)/3)^2 / dx^2 with v2,
// first-order (x - v1)^2 / dx^2 without; x = the larger root of
// sum(terms) = 1.  Where both axes enter and x is not above both axes'
// upwind values, x comes from the axis with the smaller one alone (+ 2 dx/3
// second order, + dx first order).
// The iterate as the update reads it: kBig outside the domain (a neighbour
// beyond the edge is not known).  From a global array, or from an LDS tile
// of width w whose origin is global cell (i0, j0).
struct GlobalIterate {
  const double* __restrict__ D;
  int nr, nz;
  __device__ __forceinline__ double at(int i, int j) const {
    return (i >= 0 && i < nr && j >= 0 && j < nz) ? D[(int64_t)i * nz + j] : kBig;
  }
};
struct TileIterate {
  const double* T;
  int i0, j0, w, nr, nz;
  __device__ __forceinline__ double at(int i, int j) const {
    return (i >= 0 && i < nr && j >= 0 && j < nz) ? T[(i - i0) * w + (j - j0)] : kBig;
  }
};

template <class IT>
__device__ __forceinline__ double eikonal_update(const IT& D, int i, int j, double side,
                                                 double dx) {
  const double idx2 = 1.0 / (dx * dx);
  const double aa = 9.0 / 4.0;
  double a = 0.0, b = 0.0, c = 0.0;
  double tps[2], steps[2];
  bool used[2];
#pragma unroll
  for (int axis = 0; axis < 2; ++axis) {
    double v1 = kBig, v2 = kBig;
#pragma unroll
    for (int s = -1; s <= 1; s += 2) {
      const int i1 = axis == 0 ? i + s : i, j1 = axis == 0 ? j : j + s;
      const int i2 = axis == 0 ? i + 2 * s : i, j2 = axis == 0 ? j : j + 2 * s;
      double n1 = kBig, n2 = kBig;
      const double u1 = D.at(i1, j1);
      if (u1 < kBig) n1 = side * u1;
      const double u2 = D.at(i2, j2);
      if (u2 < kBig) n2 = side * u2;
      if (n1 < v1) {
        v2 = (n2 <= n1) ? n2 : kBig;
        v1 = n1;
      }
    }
    const bool second = v1 < kBig && v2 < kBig;
    const bool first = v1 < kBig && !second;
    const double w1 = (v1 < kBig) ? v1 : 0.0;
    const double w2 = second ? v2 : 0.0;
    const double tp = second ? (4.0 * w1 - w2) / 3.0 : w1;
    a = a + (second ? idx2 * aa : (first ? idx2 : 0.0));
    b = b - (second ? idx2 * 2.0 * aa * tp : (first ? idx2 * 2.0 * w1 : 0.0));
    c = c + (second ? idx2 * aa * tp * tp : (first ? idx2 * w1 * w1 : 0.0));
    used[axis] = second || first;
    tps[axis] = used[axis] ? tp : kBig;
    steps[axis] = second ? (2.0 / 3.0) * dx : dx;
  }
  if (!(a > 0.0)) return kBig;
  c = c - 1.0;
  const double det = b * b - 4.0 * a * c;
  double x = (-b + sqrt(fmax(det, 0.0))) / (2.0 * a);
  if (used[0] && used[1] && (det < 0.0 || x < tps[0] || x < tps[1]))
    x = (tps[0] <= tps[1]) ? tps[0] + steps[0] : tps[1] + steps[1];
  return x;
}

// One Jacobi sweep: every cell that is not frozen takes its update from the
// previous iterate (values may move either way: the second-order update is
// not monotone in its neighbours).
__device__ __forceinline__ void reinit_sweep_at(const double* __restrict__ Din,
                                                double* __restrict__ Dout,
                                                const double* __restrict__ phi,
                                                const uint8_t* __restrict__ frozen, int nr,
                                                int nz, double dx, int i, int j) {
  const int64_t c = (int64_t)i * nz + j;
  const double dc = Din[c];
  if (frozen[c]) {
    Dout[c] = dc;
    return;
  }
  const double side = (phi[c] < 0.0) ? -1.0 : 1.0;
  const double m = eikonal_update(GlobalIterate{Din, nr, nz}, i, j, side, dx);
  Dout[c] = (m >= kBig) ? dc : side * m;
}

// S Jacobi sweeps of one 8 x 32 tile in LDS (temporal blocking): the tile
// and a halo of 2 S cells are loaded once; sweep s updates the tile with a
// halo of 2 (S - s) cells (every neighbour it reads is within two cells and
// was updated by the sweep before), so after nsw <= S sweeps the tile itself
// holds what nsw sweeps over the whole grid give, bit for bit.  Flags per
// cell: 1 frozen, 2 negative side, 4 outside the domain (its value stays
// kBig, as the global sweep treats a neighbour beyond the edge).
constexpr int kRtR = kReinitTileRows, kRtZ = kReinitTileCols;
template <int S>
struct ReinitTile {
  static constexpr int H = 2 * S, RR = kRtR + 2 * H, RZ = kRtZ + 2 * H, NC = RR * RZ;
  double a[NC], b[NC];
  uint8_t f[NC];
  uint8_t ca[NC], cb[NC];  // changed by the sweep before (ping-pong)
};

// A cell's update reads only its neighbours one and two cells away along
// each axis: where none of them changed in the sweep before, the update
// gives what it gave then, so the cell keeps its value without recomputing
// it (the result is the same bits; the first sweep updates every cell).
// Changes are compared bit for bit.  ch_in / ch_out carry them between
// launches.
//
// Launch k of a tile whose 3 x 3 tile neighbourhood (which holds its halo)
// changed nothing in launch k-1's last sweep changes nothing either: it
// only brings its output buffer up to date, which holds what launch k-2
// left (copy the iterate when launch k-1 changed the tile, clear the marks
// when launch k-2's last sweep left some).
template <int S>
__device__ __forceinline__ void reinit_tile_sweeps(const double* __restrict__ Din,
                                                   double* __restrict__ Dout,
                                                   const double* __restrict__ phi,
                                                   const uint8_t* __restrict__ frozen,
                                                   const uint8_t* __restrict__ ch_in,
                                                   uint8_t* __restrict__ ch_out,
                                                   uint8_t* __restrict__ tiles, int nr, int nz,
                                                   double dx, int tr, int tc, int k, int nsw,
                                                   ReinitTile<S>& L) {
  using Tl = ReinitTile<S>;
  static_assert(Tl::H <= kRtR && Tl::H <= kRtZ, "the halo must lie in the neighbour tiles");
  const int tid = threadIdx.x, nth = blockDim.x;
  const int ti0 = tr * kRtR, tj0 = tc * kRtZ;
  const int ntc = reinit_tile_cols(nz), nt = reinit_tile_count(nr, nz);
  const int ntr = nt / ntc, t = tr * ntc + tc;
  const int p = k & 1;
  uint8_t* any_w = tiles + (size_t)p * nt;
  const uint8_t* any_p = tiles + (size_t)(p ^ 1) * nt;
  uint8_t* last_w = tiles + (size_t)(2 + p) * nt;
  const uint8_t* last_p = tiles + (size_t)(2 + (p ^ 1)) * nt;
  int near = 0;
  if (tid < 9) {
    const int r = tr + tid / 3 - 1, q = tc + tid % 3 - 1;
    if (r >= 0 && r < ntr && q >= 0 && q < ntc) near = last_p[r * ntc + q];
  }
  if (!__syncthreads_or(near)) {
    const bool copy = any_p[t] != 0, clear = last_w[t] != 0;
    __syncthreads();  // every thread has read last_w[t]
    if (copy || clear)
      for (int e = tid; e < kRtR * kRtZ; e += nth) {
        const int i = ti0 + e / kRtZ, j = tj0 + e % kRtZ;
        if (i >= nr || j >= nz) continue;
        const int64_t c = (int64_t)i * nz + j;
        if (copy) Dout[c] = Din[c];
        if (clear) ch_out[c] = 0;
      }
    if (tid == 0) {
      any_w[t] = 0;
      last_w[t] = 0;
    }
    return;
  }
  __shared__ int any_change;
  if (tid == 0) any_change = 0;
  for (int e = tid; e < Tl::NC; e += nth) {
    const int i = ti0 - Tl::H + e / Tl::RZ, j = tj0 - Tl::H + e % Tl::RZ;
    if (i >= 0 && i < nr && j >= 0 && j < nz) {
      const int64_t c = (int64_t)i * nz + j;
      L.a[e] = Din[c];
      L.f[e] = uint8_t((frozen[c] ? 1 : 0) | (phi[c] < 0.0 ? 2 : 0));
      L.ca[e] = ch_in[c];
    } else {
      L.a[e] = kBig;
      L.f[e] = 4;
      L.ca[e] = 0;
    }
  }
  __syncthreads();
  double* A = L.a;
  double* B = L.b;
  uint8_t* CA = L.ca;
  uint8_t* CB = L.cb;
  for (int s = 0; s < nsw; ++s) {
    const int h = Tl::H - 2 * (s + 1);  // halo of the cells this sweep updates
    const int w = kRtZ + 2 * h, n = (kRtR + 2 * h) * w;
    const TileIterate it{A, ti0 - Tl::H, tj0 - Tl::H, Tl::RZ, nr, nz};
    for (int e = tid; e < n; e += nth) {
      const int a = e / w - h, b = e % w - h;  // tile-relative
      const int li = (a + Tl::H) * Tl::RZ + (b + Tl::H);
      const uint8_t fl = L.f[li];
      const double dc = A[li];
      double v = dc;
      const int RZ = Tl::RZ;
      const bool inputs_changed = CA[li - 2 * RZ] | CA[li - RZ] | CA[li + RZ] | CA[li + 2 * RZ] |
                                  CA[li - 2] | CA[li - 1] | CA[li + 1] | CA[li + 2];
      if (!(fl & 5) && inputs_changed) {
        const double side = (fl & 2) ? -1.0 : 1.0;
        const double m = eikonal_update(it, ti0 + a, tj0 + b, side, dx);
        v = (m >= kBig) ? dc : side * m;
      }
      B[li] = v;
      const bool ch = __double_as_longlong(v) != __double_as_longlong(dc);
      CB[li] = uint8_t(ch);
      if (ch && a >= 0 && a < kRtR && b >= 0 && b < kRtZ) any_change = 1;
    }
    __syncthreads();
    double* t = A;
    A = B;
    B = t;
    uint8_t* tc = CA;
    CA = CB;
    CB = tc;
  }
  int last = 0;
  for (int e = tid; e < kRtR * kRtZ; e += nth) {
    const int a = e / kRtZ, b = e % kRtZ;
    const int i = ti0 + a, j = tj0 + b;
    if (i < nr && j < nz) {
      const int li = (a + Tl::H) * Tl::RZ + (b + Tl::H);
      Dout[(int64_t)i * nz + j] = A[li];
      ch_out[(int64_t)i * nz + j] = CA[li];
      last |= CA[li];
    }
  }
  last = __syncthreads_or(last);
  if (tid == 0) {
    any_w[t] = uint8_t(any_change != 0);
    last_w[t] = uint8_t(last != 0);
  }
}

template <int S>
__global__ void __launch_bounds__(256) reinit_tiles_kernel(const double* __restrict__ Din,
                                                           double* __restrict__ Dout,
                                                           const double* __restrict__ phi,
                                                           const uint8_t* __restrict__ frozen,
                                                           const uint8_t* __restrict__ ch_in,
                                                           uint8_t* __restrict__ ch_out,
                                                           uint8_t* __restrict__ tiles, int nr,
                                                           int nz, double dx, int k, int nsw) {
  __shared__ ReinitTile<S> L;
  reinit_tile_sweeps<S>(Din, Dout, phi, frozen, ch_in, ch_out, tiles, nr, nz, dx, blockIdx.y,
                        blockIdx.x, k, nsw, L);
}

// Sweeps per launch of the tiled form (16 x 40 cells in LDS, 11 KB; the
// halo costs 1.34x the tile's own updates at S = 2).
constexpr int kReinitSweepsPerLaunch = 2;

__global__ void reinit_sweep_kernel(const double* __restrict__ Din, double* __restrict__ Dout,
                                    const double* __restrict__ phi,
                                    const uint8_t* __restrict__ frozen, int nr, int nz,
                                    double dx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  reinit_sweep_at(Din, Dout, phi, frozen, nr, nz, dx, i, j);
}

// Also leaves the tile flags zero for the next reinit.
__device__ __forceinline__ void reinit_final_at(const double* __restrict__ phi,
                                                const double* __restrict__ D,
                                                const double* __restrict__ dist,
                                                const uint8_t* __restrict__ frozen,
                                                uint8_t* __restrict__ tiles,
                                                double* __restrict__ out,
                                                uint8_t* __restrict__ masked, int nr, int nz,
                                                int i, int j, double narrow) {
  if (i % kReinitTileRows == 0 && j % kReinitTileCols == 0) {
    const int nt = reinit_tile_count(nr, nz);
    const int t = (i / kReinitTileRows) * reinit_tile_cols(nz) + j / kReinitTileCols;
    for (int r = 0; r < 4; ++r) tiles[(size_t)r * nt + t] = 0;
  }
  const int64_t c = (int64_t)i * nz + j;
  const double d = frozen[c] ? dist[c] : D[c];
  const bool m = !(fabs(d) <= narrow);
  out[c] = m ? phi[c] : d;
  if (masked) masked[c] = m ? 1 : 0;
}

__global__ void reinit_final_kernel(const double* __restrict__ phi, const double* __restrict__ D,
                                    ReinitWork w, double* __restrict__ out,
                                    uint8_t* __restrict__ masked, int nr, int nz, double narrow) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  reinit_final_at(phi, D, w.dist, w.frozen, w.tiles, out, masked, nr, nz, i, j, narrow);
}

// Soft batches: one row segment of 256 cells of one case per workgroup;
// phi is reinitialised in place.  A case sweeps its own count of
// iterations: sweeps beyond it exit, and its final pass reads the buffer
// its own last sweep wrote.
__global__ void soft_batch_reinit_init_kernel(const SoftCase* __restrict__ cs,
                                              const uint32_t* __restrict__ blocks) {
  const SoftBlock b = soft_block(blocks);
  const SoftCase& C = cs[b.c];
  if (b.j >= C.nz) return;
  reinit_init_at(C.phi, C.rw.d0, C.rw.d1, C.rw.dist, C.rw.frozen, C.rw.ch0, C.rw.ch1, C.rw.tiles,
                 C.nr, C.nz, C.dx, b.i, b.j);
}

// kReinitSweepsPerLaunch sweeps (fewer when the case has fewer left) of one
// 8 x 32 tile of one case (the batch's reinit tile table); launch k reads
// the first buffer when k is even.
__global__ void __launch_bounds__(256) soft_batch_reinit_tiles_kernel(
    const SoftCase* __restrict__ cs, const uint32_t* __restrict__ tiles, int q0) {
  constexpr int S = kReinitSweepsPerLaunch;
  __shared__ ReinitTile<S> L;
  const uint32_t v = tiles[blockIdx.x];
  const SoftCase& C = cs[v & 1023u];
  const int nsw = std::min(S, C.reinit_iterations - q0);
  if (nsw <= 0) return;
  const int k = q0 / S;
  const bool odd = k & 1;
  reinit_tile_sweeps<S>(odd ? C.rw.d1 : C.rw.d0, odd ? C.rw.d0 : C.rw.d1, C.phi, C.rw.frozen,
                        odd ? C.rw.ch1 : C.rw.ch0, odd ? C.rw.ch0 : C.rw.ch1, C.rw.tiles, C.nr,
                        C.nz, C.dx, int((v >> 10) & 8191u), int(v >> 23), k, nsw, L);
}

__global__ void soft_batch_reinit_final_kernel(const SoftCase* __restrict__ cs,
                                               const uint32_t* __restrict__ blocks) {
  const SoftBlock b = soft_block(blocks);
  const SoftCase& C = cs[b.c];
  if (b.j >= C.nz) return;
  constexpr int S = kReinitSweepsPerLaunch;
  const int launches = (C.reinit_iterations + S - 1) / S;
  const double* D = (launches & 1) ? C.rw.d1 : C.rw.d0;
  reinit_final_at(C.phi, D, C.rw.dist, C.rw.frozen, C.rw.tiles, C.phi, nullptr, C.nr, C.nz, b.i,
                  b.j, C.reinit_band);
}

hipError_t launch_soft_batch_reinit(const SoftCase* cases, const SoftBatchDims& d, int iterations,
                                    hipStream_t s) {
  const int g = d.ncell_blocks;
  soft_batch_reinit_init_kernel<<<g, 256, 0, s>>>(cases, d.cell_blocks);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  for (int q = 0; q < iterations; q += kReinitSweepsPerLaunch) {
    soft_batch_reinit_tiles_kernel<<<d.nreinit_tiles, 256, 0, s>>>(cases, d.reinit_tiles, q);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  soft_batch_reinit_final_kernel<<<g, 256, 0, s>>>(cases, d.cell_blocks);
  return hipGetLastError();
}

hipError_t launch_reinit(const double* phi, double* out, uint8_t* masked, int nr, int nz,
                         double dx, double narrow, int iterations, const ReinitWork& w,
                         hipStream_t s) {
  const dim3 grid(ceil_div(nz, 256), nr);
  reinit_init_kernel<<<grid, 256, 0, s>>>(phi, w, nr, nz, dx);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  double* a = w.d0;
  double* b = w.d1;
  constexpr int S = kReinitSweepsPerLaunch;
  const dim3 tiles(ceil_div(nz, kRtZ), ceil_div(nr, kRtR));
  uint8_t* ca = w.ch0;
  uint8_t* cb = w.ch1;
  for (int q = 0; q < iterations; q += S) {
    reinit_tiles_kernel<S><<<tiles, 256, 0, s>>>(a, b, phi, w.frozen, ca, cb, w.tiles, nr, nz, dx,
                                                 q / S, std::min(S, iterations - q));
    if ((e = hipGetLastError()) != hipSuccess) return e;
    double* t = a;
    a = b;
    b = t;
    uint8_t* tc = ca;
    ca = cb;
    cb = tc;
  }
  reinit_final_kernel<<<grid, 256, 0, s>>>(phi, a, w, out, masked, nr, nz, narrow);
  return hipGetLastError();
}

}  // namespace axs

using namespace axs;

extern "C" {

axs_status axs_reinit_distance_f64(const double* phi, double* out, uint8_t* masked, int nr,
                                   int nz, double dx, double narrow, int iterations,
                                   void* stream) {
  AXS_CHECK_ARG(phi && out, "null pointer");
  AXS_CHECK_ARG(nr >= 1 && nz >= 1, "empty grid");
  AXS_CHECK_ARG(dx > 0.0 && narrow > 0.0, "dx and narrow must be positive");
  if (iterations < 0) iterations = reinit_iterations(narrow, dx);
  AXS_CHECK_ARG(iterations <= 1 << 16, "too many iterations");
  const size_t n = (size_t)nr * nz;
  ReinitWork w{};
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&w.d0), sizeof(double) * n);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&w.d1), sizeof(double) * n);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&w.dist), sizeof(double) * n);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&w.frozen), n);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&w.ch0), n);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&w.ch1), n);
  const size_t tb = reinit_tile_flag_bytes(nr, nz);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&w.tiles), tb);
  hipStream_t s = as_stream(stream);
  if (e == hipSuccess) e = hipMemsetAsync(w.tiles, 0, tb, s);
  if (e == hipSuccess) e = launch_reinit(phi, out, masked, nr, nz, dx, narrow, iterations, w, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (w.d0) (void)hipFree(w.d0);
  if (w.d1) (void)hipFree(w.d1);
  if (w.dist) (void)hipFree(w.dist);
  if (w.frozen) (void)hipFree(w.frozen);
  if (w.ch0) (void)hipFree(w.ch0);
  if (w.ch1) (void)hipFree(w.ch1);
  if (w.tiles) (void)hipFree(w.tiles);
  if (e != hipSuccess) {
    set_error(std::string("reinit: ") + hipGetErrorString(e));
    return AXS_ERR_HIP;
  }
  return AXS_OK;
}

}  // extern "C"
