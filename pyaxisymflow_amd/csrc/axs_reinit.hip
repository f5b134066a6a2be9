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

// The sweeps run on 8 x 32 tiles (kReinitTileRows x kReinitTileCols), S
// sweeps per launch; a tile's sweeps read its cells and a halo of 2 S
// (its read region).
constexpr int kRtR = kReinitTileRows, kRtZ = kReinitTileCols;
constexpr int kReinitSweepsPerLaunch = 2;
constexpr int kReinitHalo = 2 * kReinitSweepsPerLaunch;
static_assert(kReinitHalo <= kRtR && kReinitHalo <= kRtZ,
              "a read region must lie in the 3 x 3 tiles around its tile");
__host__ __device__ __forceinline__ int reinit_tile_cols(int nz) {
  return (nz + kRtZ - 1) / kRtZ;
}
__host__ __device__ __forceinline__ int reinit_tile_count(int nr, int nz) {
  return ((nr + kRtR - 1) / kRtR) * reinit_tile_cols(nz);
}

int reinit_tile_total(int nr, int nz) { return reinit_tile_count(nr, nz); }

// Work lists.  A launch takes only the tiles on its list: those with a
// cell changed by the launch before's last sweep in their read region
// (flag bit 0: run the sweeps), and those that ran in the launch before
// (bit 1: one that does not run again copies its iterate into the other
// buffer and clears its marks there, so that a tile off the lists has the
// same iterate in both buffers and no marks).  Every other tile would
// compute what it holds.  Flags (ReinitWork::tiles): one word per tile and
// launch parity, set with atomicOr by the tiles that list it (the one that
// finds it zero appends the tile) and cleared when the tile is taken.
// Lists: two of `cap` entries (by parity); counters: three, launch k reads
// k % 3, fills (k + 1) % 3 and clears (k + 2) % 3.
//
// Both iterate buffers start as the initial iterate (kBig off the
// interface), with the interface cells marked changed in the first and
// nothing in the second: a cell none of whose neighbours is an interface
// cell has only kBig neighbours, which leave it kBig, so marking the
// interface cells alone starts the sweeps as marking every cell would.
// The first list holds the tiles with an interface cell in their read
// region.

// floor(x / d) for d > 0 and any x
__device__ __forceinline__ int floor_div(int x, int d) { return (x >= 0) ? x / d : -((d - 1 - x) / d); }
__device__ __forceinline__ void reinit_init_at(const double* __restrict__ phi,
                                               double* __restrict__ D, double* __restrict__ D1,
                                               double* __restrict__ dist,
                                               uint8_t* __restrict__ frozen,
                                               uint8_t* __restrict__ ch0, uint8_t* __restrict__ ch1,
                                               unsigned* __restrict__ tiles, int nr, int nz,
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
  if (fr) {  // the tiles whose read region holds the cell run the first launch
    const int ntc = reinit_tile_cols(nz), ntr = reinit_tile_count(nr, nz) / ntc;
    const int r0 = max(0, floor_div(i - kReinitHalo, kRtR));
    const int r1 = min(ntr - 1, floor_div(i + kReinitHalo, kRtR));
    const int q0 = max(0, floor_div(j - kReinitHalo, kRtZ));
    const int q1 = min(ntc - 1, floor_div(j + kReinitHalo, kRtZ));
    for (int r = r0; r <= r1; ++r)
      for (int q = q0; q <= q1; ++q) tiles[r * ntc + q] = 1u;
  }
}

__global__ void reinit_init_kernel(const double* __restrict__ phi, ReinitWork w, int nr, int nz,
                                   double dx) {
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 2) w.cnt[threadIdx.x] = 0u;
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
template <int S>
struct ReinitTile {
  static constexpr int H = 2 * S, RR = kRtR + 2 * H, RZ = kRtZ + 2 * H, NC = RR * RZ;
  double a[NC], b[NC];
  uint8_t f[NC];
  uint8_t ca[NC], cb[NC];  // changed by the sweep before (ping-pong)
};

// The tiles (neighbour index (dr + 1) * 3 + dc + 1, itself = 4) whose read
// region holds cell (a, b) of a tile.
__device__ __forceinline__ unsigned reinit_touch(int a, int b) {
  const unsigned rows = 2u | (a < kReinitHalo ? 1u : 0u) | (a >= kRtR - kReinitHalo ? 4u : 0u);
  const unsigned cols = 2u | (b < kReinitHalo ? 1u : 0u) | (b >= kRtZ - kReinitHalo ? 4u : 0u);
  unsigned m = 0u;
#pragma unroll
  for (int r = 0; r < 3; ++r)
    if ((rows >> r) & 1u) m |= cols << (3 * r);
  return m;
}

// Launch k of tile (tr, tc) taken from its list.  A cell's update reads
// only its neighbours one and two cells away along each axis: where none
// of them changed in the sweep before, the update gives what it gave then,
// so the cell keeps its value without recomputing it (the same bits).
// Changes are compared bit for bit; ch_in / ch_out carry them between
// launches.  flags_k / flags_n: this launch's and the next one's tile
// flags; list_n, cnt_n: the next launch's list (base + tile index) when
// `more`.
template <int S>
__device__ __forceinline__ void reinit_tile_step(
    const double* __restrict__ Din, double* __restrict__ Dout, const double* __restrict__ phi,
    const uint8_t* __restrict__ frozen, const uint8_t* __restrict__ ch_in,
    uint8_t* __restrict__ ch_out, unsigned* __restrict__ flags_k, unsigned* __restrict__ flags_n,
    unsigned* __restrict__ list_n, unsigned* __restrict__ cnt_n, unsigned base, bool more, int nr,
    int nz, double dx, int tr, int tc, int nsw, ReinitTile<S>& L) {
  using Tl = ReinitTile<S>;
  static_assert(Tl::H == kReinitHalo, "read region");
  const int tid = threadIdx.x, nth = blockDim.x;
  const int ti0 = tr * kRtR, tj0 = tc * kRtZ;
  const int ntc = reinit_tile_cols(nz), ntr = reinit_tile_count(nr, nz) / ntc;
  const int t = tr * ntc + tc;
  __shared__ unsigned s_flag, s_touch;
  if (tid == 0) {
    s_flag = flags_k[t];
    flags_k[t] = 0u;
    s_touch = 0u;
  }
  __syncthreads();
  if (!(s_flag & 1u)) {  // ran last launch, quiet since: bring the other buffer up to date
    for (int e = tid; e < kRtR * kRtZ; e += nth) {
      const int i = ti0 + e / kRtZ, j = tj0 + e % kRtZ;
      if (i >= nr || j >= nz) continue;
      const int64_t c = (int64_t)i * nz + j;
      Dout[c] = Din[c];
      ch_out[c] = 0;
    }
    __syncthreads();  // the shared words are read before the next tile sets them
    return;
  }
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
      CB[li] = uint8_t(__double_as_longlong(v) != __double_as_longlong(dc));
    }
    __syncthreads();
    double* tp = A;
    A = B;
    B = tp;
    uint8_t* tcp = CA;
    CA = CB;
    CB = tcp;
  }
  unsigned touch = 0u;
  for (int e = tid; e < kRtR * kRtZ; e += nth) {
    const int a = e / kRtZ, b = e % kRtZ;
    const int i = ti0 + a, j = tj0 + b;
    if (i < nr && j < nz) {
      const int li = (a + Tl::H) * Tl::RZ + (b + Tl::H);
      Dout[(int64_t)i * nz + j] = A[li];
      ch_out[(int64_t)i * nz + j] = CA[li];
      if (CA[li]) touch |= reinit_touch(a, b);
    }
  }
  if (touch) atomicOr(&s_touch, touch);
  __syncthreads();
  if (more && tid < 9) {
    const int r = tr + tid / 3 - 1, q = tc + tid % 3 - 1;
    const unsigned bits = ((s_touch >> tid) & 1u) | (tid == 4 ? 2u : 0u);
    if (bits && r >= 0 && r < ntr && q >= 0 && q < ntc) {
      const int u = r * ntc + q;
      if (atomicOr(&flags_n[u], bits) == 0u) list_n[atomicAdd(cnt_n, 1u)] = base + (unsigned)u;
    }
  }
  __syncthreads();  // the shared words and L are read before the next tile
}

// The first list: the tiles with a flag set by the init pass.
__global__ void reinit_first_list_kernel(const unsigned* __restrict__ flags, int nt,
                                         unsigned* __restrict__ list, unsigned* __restrict__ cnt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  list_append(t < nt && flags[t] != 0u, t, reinterpret_cast<int*>(list), cnt);
}

// Launch k over its list: every workgroup takes entries blockIdx.x,
// blockIdx.x + gridDim.x, ...
template <int S>
__global__ void __launch_bounds__(256) reinit_list_kernel(
    const double* __restrict__ Din, double* __restrict__ Dout, const double* __restrict__ phi,
    const uint8_t* __restrict__ frozen, const uint8_t* __restrict__ ch_in,
    uint8_t* __restrict__ ch_out, unsigned* __restrict__ flags, unsigned* __restrict__ list,
    unsigned* __restrict__ cnt, int nr, int nz, double dx, int k, int nsw, int more) {
  __shared__ ReinitTile<S> L;
  const int nt = reinit_tile_count(nr, nz), ntc = reinit_tile_cols(nz);
  const int p = k & 1;
  if (blockIdx.x == 0 && threadIdx.x == 0) cnt[(k + 2) % 3] = 0u;
  const int n = (int)cnt[k % 3];
  for (int idx = blockIdx.x; idx < n; idx += gridDim.x) {
    const int t = (int)list[(size_t)p * nt + idx];
    reinit_tile_step<S>(Din, Dout, phi, frozen, ch_in, ch_out, flags + (size_t)p * nt,
                        flags + (size_t)(p ^ 1) * nt, list + (size_t)(p ^ 1) * nt,
                        cnt + (k + 1) % 3, 0u, more != 0, nr, nz, dx, t / ntc, t % ntc, nsw, L);
  }
}

// Also leaves the tile flags zero for the next reinit.
__device__ __forceinline__ void reinit_final_at(const double* __restrict__ phi,
                                                const double* __restrict__ D,
                                                const double* __restrict__ dist,
                                                const uint8_t* __restrict__ frozen,
                                                unsigned* __restrict__ tiles,
                                                double* __restrict__ out,
                                                uint8_t* __restrict__ masked, int nr, int nz,
                                                int i, int j, double narrow) {
  if (i % kRtR == 0 && j % kRtZ == 0) {
    const int nt = reinit_tile_count(nr, nz);
    const int t = (i / kRtR) * reinit_tile_cols(nz) + j / kRtZ;
    tiles[t] = 0u;
    tiles[(size_t)nt + t] = 0u;
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
// iterations: its tiles leave the lists after its last launch, and its
// final pass reads the buffer its own last launch wrote.  The lists are
// the batch's, over its reinit tile table (entry = the case's first table
// index + the tile's index in the case).
template <typename B>
__global__ void soft_batch_reinit_init_kernel(const SoftCaseT<B>* __restrict__ cs,
                                              const uint32_t* __restrict__ blocks,
                                              unsigned* __restrict__ cnt) {
  if (blockIdx.x == 0 && threadIdx.x < 2) cnt[threadIdx.x] = 0u;
  const SoftBlock b = soft_block(blocks);
  const SoftCaseT<B>& C = cs[b.c];
  if (b.j >= C.nz) return;
  reinit_init_at(C.phi, C.rw.d0, C.rw.d1, C.rw.dist, C.rw.frozen, C.rw.ch0, C.rw.ch1, C.rw.tiles,
                 C.nr, C.nz, C.dx, b.i, b.j);
}

template <typename B>
__global__ void soft_batch_reinit_first_list_kernel(const SoftCaseT<B>* __restrict__ cs,
                                                    const uint32_t* __restrict__ tiles, int n,
                                                    unsigned* __restrict__ list,
                                                    unsigned* __restrict__ cnt) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  bool take = false;
  if (idx < n) {
    const uint32_t v = tiles[idx];
    const SoftCaseT<B>& C = cs[v & 1023u];
    const int t = int((v >> 10) & 8191u) * reinit_tile_cols(C.nz) + int(v >> 23);
    take = C.reinit_iterations > 0 && C.rw.tiles[t] != 0u;
  }
  list_append(take, idx, reinterpret_cast<int*>(list), cnt);
}

template <typename B>
__global__ void __launch_bounds__(256) soft_batch_reinit_list_kernel(
    const SoftCaseT<B>* __restrict__ cs, const uint32_t* __restrict__ tiles, int ntab,
    unsigned* __restrict__ list, unsigned* __restrict__ cnt, int k) {
  constexpr int S = kReinitSweepsPerLaunch;
  __shared__ ReinitTile<S> L;
  const int p = k & 1;
  if (blockIdx.x == 0 && threadIdx.x == 0) cnt[(k + 2) % 3] = 0u;
  const int n = (int)cnt[k % 3];
  for (int idx = blockIdx.x; idx < n; idx += gridDim.x) {
    const unsigned e = list[(size_t)p * ntab + idx];
    const uint32_t v = tiles[e];
    const SoftCaseT<B>& C = cs[v & 1023u];
    const int nsw = std::min(S, C.reinit_iterations - k * S);
    if (nsw <= 0) continue;  // block-uniform
    const int tr = int((v >> 10) & 8191u), tc = int(v >> 23);
    const int nt = reinit_tile_count(C.nr, C.nz);
    const bool odd = p;
    reinit_tile_step<S>(odd ? C.rw.d1 : C.rw.d0, odd ? C.rw.d0 : C.rw.d1, C.phi, C.rw.frozen,
                        odd ? C.rw.ch1 : C.rw.ch0, odd ? C.rw.ch0 : C.rw.ch1,
                        C.rw.tiles + (size_t)p * nt, C.rw.tiles + (size_t)(p ^ 1) * nt,
                        list + (size_t)(p ^ 1) * ntab, cnt + (k + 1) % 3,
                        (unsigned)C.reinit_tile_base, (k + 1) * S < C.reinit_iterations, C.nr,
                        C.nz, C.dx, tr, tc, nsw, L);
  }
}

template <typename B>
__global__ void soft_batch_reinit_final_kernel(const SoftCaseT<B>* __restrict__ cs,
                                               const uint32_t* __restrict__ blocks) {
  const SoftBlock b = soft_block(blocks);
  const SoftCaseT<B>& C = cs[b.c];
  if (b.j >= C.nz) return;
  constexpr int S = kReinitSweepsPerLaunch;
  const int launches = (C.reinit_iterations + S - 1) / S;
  const double* D = (launches & 1) ? C.rw.d1 : C.rw.d0;
  reinit_final_at(C.phi, D, C.rw.dist, C.rw.frozen, C.rw.tiles, C.phi, nullptr, C.nr, C.nz, b.i,
                  b.j, C.reinit_band);
}

// Workgroups of a list launch: at most one per entry, and at most 1536 (six
// per CU of the MI355X's 256); more entries loop.  Measured on a 24-case soft
// batch: 1536 workgroups beat 2048 by 6-7% per step (the extra ones cost
// their dispatch even when the list has shrunk to a few tiles) and 1024 or
// 1280 by 1-2%; the resident count the occupancy query reports did not do
// as well.
static int reinit_list_grid(int entries) { return std::max(1, std::min(entries, 1536)); }

template <typename B>
hipError_t launch_soft_batch_reinit(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                    int iterations, hipStream_t s) {
  const int g = d.ncell_blocks;
  soft_batch_reinit_init_kernel<B><<<g, 256, 0, s>>>(cases, d.cell_blocks, d.reinit_cnt);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  soft_batch_reinit_first_list_kernel<B><<<ceil_div(d.nreinit_tiles, 256), 256, 0, s>>>(
      cases, d.reinit_tiles, d.nreinit_tiles, d.reinit_list, d.reinit_cnt);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const int grid = reinit_list_grid(d.nreinit_tiles);
  for (int k = 0; k * kReinitSweepsPerLaunch < iterations; ++k) {
    soft_batch_reinit_list_kernel<B><<<grid, 256, 0, s>>>(cases, d.reinit_tiles, d.nreinit_tiles,
                                                       d.reinit_list, d.reinit_cnt, k);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  soft_batch_reinit_final_kernel<B><<<g, 256, 0, s>>>(cases, d.cell_blocks);
  return hipGetLastError();
}

template hipError_t launch_soft_batch_reinit<double>(const SoftCaseT<double>*,
                                                     const SoftBatchDims&, int, hipStream_t);
template hipError_t launch_soft_batch_reinit<float>(const SoftCaseT<float>*,
                                                    const SoftBatchDims&, int, hipStream_t);

hipError_t launch_reinit(const double* phi, double* out, uint8_t* masked, int nr, int nz,
                         double dx, double narrow, int iterations, const ReinitWork& w,
                         hipStream_t s) {
  const dim3 grid(ceil_div(nz, 256), nr);
  reinit_init_kernel<<<grid, 256, 0, s>>>(phi, w, nr, nz, dx);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int nt = reinit_tile_count(nr, nz);
  reinit_first_list_kernel<<<ceil_div(nt, 256), 256, 0, s>>>(w.tiles, nt, w.list, w.cnt);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  double* a = w.d0;
  double* b = w.d1;
  constexpr int S = kReinitSweepsPerLaunch;
  uint8_t* ca = w.ch0;
  uint8_t* cb = w.ch1;
  const int lg = reinit_list_grid(nt);
  for (int q = 0, k = 0; q < iterations; q += S, ++k) {
    reinit_list_kernel<S><<<lg, 256, 0, s>>>(a, b, phi, w.frozen, ca, cb, w.tiles, w.list, w.cnt,
                                             nr, nz, dx, k, std::min(S, iterations - q),
                                             q + S < iterations ? 1 : 0);
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
  const size_t nt = (size_t)reinit_tile_total(nr, nz);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&w.tiles), 2 * nt * sizeof(unsigned));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&w.list), 2 * nt * sizeof(unsigned));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&w.cnt), 3 * sizeof(unsigned));
  hipStream_t s = as_stream(stream);
  if (e == hipSuccess) e = hipMemsetAsync(w.tiles, 0, 2 * nt * sizeof(unsigned), s);
  if (e == hipSuccess) e = launch_reinit(phi, out, masked, nr, nz, dx, narrow, iterations, w, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (w.d0) (void)hipFree(w.d0);
  if (w.d1) (void)hipFree(w.d1);
  if (w.dist) (void)hipFree(w.dist);
  if (w.frozen) (void)hipFree(w.frozen);
  if (w.ch0) (void)hipFree(w.ch0);
  if (w.ch1) (void)hipFree(w.ch1);
  if (w.tiles) (void)hipFree(w.tiles);
  if (w.list) (void)hipFree(w.list);
  if (w.cnt) (void)hipFree(w.cnt);
  if (e != hipSuccess) {
    set_error(std::string("reinit: ") + hipGetErrorString(e));
    return AXS_ERR_HIP;
  }
  return AXS_OK;
}

}  // extern "C"
