// Least-squares extrapolation of the reference map (row A14):
// core/src/extrapolate_using_least_squares.hpp:142-446 as instantiated by
// extrapolate_using_least_squares_till_first_order<short, double> (:450-467).
//
// Each sweep of the reference scans every cell, collects the cells with
// current ^ target set and at least one known 3x3 neighbour, fits
// eta = c0 + c1 x + c2 y on a shifted 3x3 window of known cells and only
// then marks the collected cells known (:441-444).  A sweep is therefore
// data-parallel: here one thread per cell fits into side buffers, and a
// second pass commits values and flags.  The arithmetic follows the
// reference's order with FMA contraction off, so results match it bit for
// bit.  Out-of-range cells read as flag 0 (the reference has no bounds
// checks there, :205-219, 276-280).
#include "axs_lsq.hpp"
#include "axs_softbatch.hpp"

#include <vector>

namespace axs {

__device__ __forceinline__ short lsq_flag(const int16_t* f, int n0, int n1, int j, int i) {
  return (j < 0 || j >= n0 || i < 0 || i >= n1) ? short(0) : short(f[(int64_t)j * n1 + i]);
}

// ExtrapolationConfig.hpp:17-19, 131-158 (grid size 3: start = code - 3).
__device__ __forceinline__ short lsq_start_code(short positional_sum, short total_sum) {
  float temp = (positional_sum > short(4)) ? float(ceil(0.5 * float(positional_sum)) * 2.0)
                                           : float(positional_sum);
  short code = short(2.0 * double(temp) / double(total_sum));
  return short(short(2 * code - short(6)) / short(2));
}

// least_squares.hpp:10-59 for 3 coefficients and 2 right-hand sides.
__device__ void lsq_gauss_3x2(double A[3][5], double sol[2][3]) {
  for (int i = 0; i < 3; i++) {
    int max_row = i;
    double max_element = fabs(A[i][i]);
    for (int k = i + 1; k < 3; k++) {
      if (fabs(A[k][i]) > max_element) {
        max_element = fabs(A[k][i]);
        max_row = k;
      }
    }
    for (int k = i; k < 5; k++) {
      double tmp = A[max_row][k];
      A[max_row][k] = A[i][k];
      A[i][k] = tmp;
    }
    for (int k = i + 1; k < 3; k++) {
      double c = -A[k][i] / A[i][i];
      A[k][i] = 0.0;
      for (int j = i + 1; j < 5; j++) A[k][j] += c * A[i][j];
    }
  }
  for (int comp = 0; comp < 2; ++comp) {
    for (int i = 2; i >= 0; --i) {
      sol[comp][i] = A[i][3 + comp] / A[i][i];
      for (int k = i - 1; k >= 0; --k) A[k][3 + comp] -= A[k][i] * sol[comp][i];
    }
  }
}

// Fit of cell (j, i) from the known cells around it
// (extrapolate_using_least_squares.hpp:190-305); false when no 3x3
// neighbour is known (the cell waits for a later sweep).
__device__ __forceinline__ bool lsq_fit_cell(const int16_t* __restrict__ cur,
                                             const double* __restrict__ ex,
                                             const double* __restrict__ ey,
                                             const double* __restrict__ gx,
                                             const double* __restrict__ gy, int n0, int n1,
                                             int j, int i, double& vx, double& vy) {
  short local_sum = 0, lx = 0, ly = 0;
  for (int jl = -1; jl <= 1; ++jl)
    for (int il = -1; il <= 1; ++il) {
      const short f = lsq_flag(cur, n0, n1, j + jl, i + il);
      local_sum += (jl == 0 && il == 0) ? short(0) : f;
      lx += f * short(il + 1);
      ly += f * short(jl + 1);
    }
  if (!local_sum) return false;
  const int s_x = i + lsq_start_code(lx, local_sum);
  const int s_y = j + lsq_start_code(ly, local_sum);
  double L0[9], L1[9], L2[9], r0[9], r1[9];
  int n = 0;
  for (int jl = s_y; jl < s_y + 3; ++jl) {
    for (int il = s_x; il < s_x + 3; ++il, ++n) {
      const bool in = (jl >= 0 && jl < n0 && il >= 0 && il < n1);
      const double m = in ? double(cur[(int64_t)jl * n1 + il]) : 0.0;
      const double xv = in ? gx[il] : 0.0;
      const double yv = in ? gy[jl] : 0.0;
      L0[n] = m;
      L1[n] = m * xv;
      L2[n] = m * yv;
      const int64_t g = (int64_t)jl * n1 + il;
      r0[n] = m * (in ? ex[g] : 0.0);
      r1[n] = m * (in ? ey[g] : 0.0);
    }
  }
  const double* Ls[3] = {L0, L1, L2};
  double B[3][5];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double s = 0.0;
      for (int q = 0; q < 9; ++q) s += Ls[a][q] * Ls[b][q];
      B[a][b] = s;
    }
  for (int a = 0; a < 3; ++a) {
    double s = 0.0;
    for (int q = 0; q < 9; ++q) s += Ls[a][q] * r0[q];
    B[a][3] = s;
    s = 0.0;
    for (int q = 0; q < 9; ++q) s += Ls[a][q] * r1[q];
    B[a][4] = s;
  }
  double sol[2][3];
  lsq_gauss_3x2(B, sol);
  const double px = gx[i], py = gy[j];
  const double E[3] = {1.0, 1.0 * px, 1.0 * py};
  double e = 0.0;
  for (int q = 0; q < 3; ++q) e += sol[0][q] * E[q];
  vx = e;
  e = 0.0;
  for (int q = 0; q < 3; ++q) e += sol[1][q] * E[q];
  vy = e;
  return true;
}

// One sweep, phase 1: select + fit.  pend[c] = 1 when cell c is extrapolated
// this sweep (new values in nx/ny), else 0.
__global__ void lsq_fit_kernel(const int16_t* __restrict__ cur, const int16_t* __restrict__ tgt,
                               const double* __restrict__ ex, const double* __restrict__ ey,
                               const double* __restrict__ gx, const double* __restrict__ gy,
                               int n0, int n1, double* __restrict__ nx, double* __restrict__ ny,
                               int8_t* __restrict__ pend, unsigned* __restrict__ counter,
                               const unsigned* __restrict__ prev_counter) {
  // A sweep after one that fitted nothing does nothing (the reference loop
  // has ended); its commit sees a zero counter and exits too.
  if (prev_counter && *prev_counter == 0u) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // column (x)
  const int j = blockIdx.y;                              // row (y)
  if (i >= n1) return;
  const int64_t c = (int64_t)j * n1 + i;
  pend[c] = 0;
  if (!(cur[c] ^ tgt[c])) return;
  double vx, vy;
  if (!lsq_fit_cell(cur, ex, ey, gx, gy, n0, n1, j, i, vx, vy)) return;
  nx[c] = vx;
  ny[c] = vy;
  pend[c] = 1;
  atomicAdd(counter, 1u);
}

// The same sweep over a list of cells (the ones with current ^ target set
// when the list was made; entries that became known are skipped by the same
// test).  pend is indexed by list entry.
__global__ void lsq_fit_list_kernel(const int16_t* __restrict__ cur,
                                    const int16_t* __restrict__ tgt,
                                    const double* __restrict__ ex, const double* __restrict__ ey,
                                    const double* __restrict__ gx, const double* __restrict__ gy,
                                    int n0, int n1, const int* __restrict__ list,
                                    const unsigned* __restrict__ count, double* __restrict__ nx,
                                    double* __restrict__ ny, int8_t* __restrict__ pend,
                                    unsigned* __restrict__ counter,
                                    const unsigned* __restrict__ prev_counter) {
  if (prev_counter && *prev_counter == 0u) return;
  const int nl = (int)*count;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nl; k += gridDim.x * blockDim.x) {
    const int c = list[k];
    int8_t p = 0;
    if (cur[c] ^ tgt[c]) {
      double vx, vy;
      if (lsq_fit_cell(cur, ex, ey, gx, gy, n0, n1, c / n1, c % n1, vx, vy)) {
        nx[c] = vx;
        ny[c] = vy;
        p = 1;
        atomicAdd(counter, 1u);
      }
    }
    pend[k] = p;
  }
}

// Phase 2: commit values and flags of the cells fitted in phase 1.
__global__ void lsq_commit_kernel(int16_t* __restrict__ cur, double* __restrict__ ex,
                                  double* __restrict__ ey, const double* __restrict__ nx,
                                  const double* __restrict__ ny, const int8_t* __restrict__ pend,
                                  int64_t n, const unsigned* __restrict__ counter) {
  if (*counter == 0u) return;  // nothing fitted (or the loop has ended)
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n || !pend[c]) return;
  ex[c] = nx[c];
  ey[c] = ny[c];
  cur[c] = 1;
}

__global__ void lsq_commit_list_kernel(int16_t* __restrict__ cur, double* __restrict__ ex,
                                       double* __restrict__ ey, const double* __restrict__ nx,
                                       const double* __restrict__ ny,
                                       const int8_t* __restrict__ pend,
                                       const int* __restrict__ list,
                                       const unsigned* __restrict__ count,
                                       const unsigned* __restrict__ counter) {
  if (*counter == 0u) return;
  const int nl = (int)*count;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nl; k += gridDim.x * blockDim.x) {
    if (!pend[k]) continue;
    const int c = list[k];
    ex[c] = nx[c];
    ey[c] = ny[c];
    cur[c] = 1;
  }
}

// Cells with current ^ target set, in any order (the set is what matters:
// a sweep's fits read only the previous sweep's state); *count must be zero.
__global__ void lsq_list_kernel(const int16_t* __restrict__ cur, const int16_t* __restrict__ tgt,
                                int64_t n, int* __restrict__ list, unsigned* __restrict__ count) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  list_append(c < n && (cur[c] ^ tgt[c]), (int)c, list, count);
}

__global__ void lsq_sweep_pingpong_kernel(
    const int16_t* __restrict__ cur_in, const double* __restrict__ ex_in,
    const double* __restrict__ ey_in, int16_t* __restrict__ cur_out, double* __restrict__ ex_out,
    double* __restrict__ ey_out, const int16_t* __restrict__ tgt, const double* __restrict__ gx,
    const double* __restrict__ gy, int n0, int n1, const int* __restrict__ list,
    const unsigned* __restrict__ count, unsigned* __restrict__ counter,
    const unsigned* __restrict__ prev_counter) {
  if (prev_counter && *prev_counter == 0u) return;  // the reference loop has ended
  const int nl = (int)*count;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nl; k += gridDim.x * blockDim.x) {
    const int c = list[k];
    double vx, vy;
    if ((cur_in[c] ^ tgt[c]) &&
        lsq_fit_cell(cur_in, ex_in, ey_in, gx, gy, n0, n1, c / n1, c % n1, vx, vy)) {
      ex_out[c] = vx;
      ey_out[c] = vy;
      cur_out[c] = 1;
      atomicAdd(counter, 1u);
    } else {
      ex_out[c] = ex_in[c];
      ey_out[c] = ey_in[c];
      cur_out[c] = cur_in[c];
    }
  }
}

__device__ void lsq_sweep_tail_body(
    int16_t* __restrict__ cur_a, double* __restrict__ ex_a, double* __restrict__ ey_a,
    int16_t* __restrict__ cur_b, double* __restrict__ ex_b, double* __restrict__ ey_b,
    const int16_t* __restrict__ tgt, const double* __restrict__ gx, const double* __restrict__ gy,
    int n0, int n1, const int* __restrict__ list, const unsigned* __restrict__ count,
    unsigned* __restrict__ counters, int sweeps, int* __restrict__ tail, int cap) {
  // unknown cells (u) and cells fitted by the previous sweep (f), read from
  // index `par` and written to index par ^ 1 of each
  int* ul[2] = {tail, tail + cap};
  int* fl[2] = {tail + 2 * (int64_t)cap, tail + 3 * (int64_t)cap};
  __shared__ unsigned nu[2], nf[2];
  if (counters[sweeps - 1] == 0u) {  // the loop ended within the queued sweeps
    if (threadIdx.x == 0) counters[sweeps] = 0u;
    return;
  }
  // sweep sweeps-1 wrote the second set when sweeps-1 is even: the result
  // must end there
  const bool home_b = ((sweeps - 1) & 1) == 0;
  int16_t* ci = home_b ? cur_b : cur_a;
  double* xi = home_b ? ex_b : ex_a;
  double* yi = home_b ? ey_b : ey_a;
  int16_t* co = home_b ? cur_a : cur_b;
  double* xo = home_b ? ex_a : ex_b;
  double* yo = home_b ? ey_a : ey_b;
  const int nl = (int)*count;
  unsigned ran = 0;
  for (;;) {
    const int par = ran & 1, nxt = par ^ 1;
    if (threadIdx.x == 0) {
      nu[nxt] = 0u;
      nf[nxt] = 0u;
    }
    __syncthreads();
    if (ran == 0) {
      // the whole list: fit, or copy through (the two sets differ on the
      // cells the last queued sweep fitted)
      for (int k = threadIdx.x; k < nl; k += blockDim.x) {
        const int c = list[k];
        double vx, vy;
        const bool unknown = ci[c] ^ tgt[c];
        if (unknown && lsq_fit_cell(ci, xi, yi, gx, gy, n0, n1, c / n1, c % n1, vx, vy)) {
          xo[c] = vx;
          yo[c] = vy;
          co[c] = 1;
          fl[nxt][atomicAdd(&nf[nxt], 1u)] = c;
        } else {
          xo[c] = xi[c];
          yo[c] = yi[c];
          co[c] = ci[c];
          if (unknown) ul[nxt][atomicAdd(&nu[nxt], 1u)] = c;
        }
      }
    } else {
      // cells fitted by the previous sweep: into this sweep's output set
      const int nfp = (int)nf[par];
      for (int k = threadIdx.x; k < nfp; k += blockDim.x) {
        const int c = fl[par][k];
        xo[c] = xi[c];
        yo[c] = yi[c];
        co[c] = ci[c];
      }
      // cells still unknown (equal, flag 0, in both sets): fit or keep
      const int nup = (int)nu[par];
      for (int k = threadIdx.x; k < nup; k += blockDim.x) {
        const int c = ul[par][k];
        double vx, vy;
        if (lsq_fit_cell(ci, xi, yi, gx, gy, n0, n1, c / n1, c % n1, vx, vy)) {
          xo[c] = vx;
          yo[c] = vy;
          co[c] = 1;
          fl[nxt][atomicAdd(&nf[nxt], 1u)] = c;
        } else {
          ul[nxt][atomicAdd(&nu[nxt], 1u)] = c;
        }
      }
    }
    __syncthreads();
    ++ran;
    int16_t* tc = ci;
    ci = co;
    co = tc;
    double* t = xi;
    xi = xo;
    xo = t;
    t = yi;
    yi = yo;
    yo = t;
    // the same values for every thread; with current flags inside target
    // ones every sweep but the last fills a listed cell, so the second
    // test only bounds the loop for flags that break that rule
    if (nf[nxt] == 0u || ran > (unsigned)nl) break;
  }
  // the last sweep fitted nothing and copied the previous sweep's fits
  // across: both sets now hold every fit, the home set included
  if (threadIdx.x == 0) counters[sweeps] = ran;
}

__global__ void __launch_bounds__(512) lsq_sweep_tail_kernel(
    int16_t* __restrict__ cur_a, double* __restrict__ ex_a, double* __restrict__ ey_a,
    int16_t* __restrict__ cur_b, double* __restrict__ ex_b, double* __restrict__ ey_b,
    const int16_t* __restrict__ tgt, const double* __restrict__ gx, const double* __restrict__ gy,
    int n0, int n1, const int* __restrict__ list, const unsigned* __restrict__ count,
    unsigned* __restrict__ counters, int sweeps, int* __restrict__ tail, int cap) {
  lsq_sweep_tail_body(cur_a, ex_a, ey_a, cur_b, ex_b, ey_b, tgt, gx, gy, n0, n1, list, count,
                      counters, sweeps, tail, cap);
}

// ------------------------------------------------------------ soft batches
// The doubled (2 nr, nz) grid of every case, z serving as both axes'
// coordinates as in the single-case step; each case has its own list,
// counters and tail lists.

template <typename B>
__global__ void soft_batch_lsq_list_kernel(const SoftCaseT<B>* __restrict__ cs,
                                           const uint32_t* __restrict__ blocks) {
  const SoftBlock b = soft_block(blocks);  // a row of the doubled grid
  const SoftCaseT<B>& C = cs[b.c];
  // the doubled row's physical row (both halves mirror it); outside the
  // step's zone every flag is 0 (the whole workgroup leaves together)
  if (!zone_row(C, b.i >= C.nr ? b.i - C.nr : C.nr - 1 - b.i)) return;
  const int j = b.j;
  const int64_t c = (int64_t)b.i * C.nz + j;
  list_append(j < C.nz && (C.cur[c] ^ C.tgt[c]), (int)c, C.lw.list, C.lw.count);
}

// Ping-pong sweep q over each case's list (blockIdx.y = case): sweep q reads
// the first set (cur, e1d, e2d) when q is even.
template <typename B>
__global__ void soft_batch_lsq_sweep_kernel(const SoftCaseT<B>* __restrict__ cs, int q) {
  const SoftCaseT<B>& C = cs[blockIdx.y];
  const LsqWork& w = C.lw;
  if (q > 0 && w.counters[q - 1] == 0u) return;  // the reference loop has ended
  const bool odd = q & 1;
  const int16_t* cur_in = odd ? w.cur2 : C.cur;
  const double* ex_in = odd ? w.ex2 : C.e1d;
  const double* ey_in = odd ? w.ey2 : C.e2d;
  int16_t* cur_out = odd ? C.cur : w.cur2;
  double* ex_out = odd ? C.e1d : w.ex2;
  double* ey_out = odd ? C.e2d : w.ey2;
  const int n0 = 2 * C.nr, n1 = C.nz;
  const int nl = (int)*w.count;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nl; k += gridDim.x * blockDim.x) {
    const int c = w.list[k];
    double vx, vy;
    if ((cur_in[c] ^ C.tgt[c]) &&
        lsq_fit_cell(cur_in, ex_in, ey_in, C.z, C.z, n0, n1, c / n1, c % n1, vx, vy)) {
      ex_out[c] = vx;
      ey_out[c] = vy;
      cur_out[c] = 1;
      atomicAdd(w.counters + q, 1u);
    } else {
      ex_out[c] = ex_in[c];
      ey_out[c] = ey_in[c];
      cur_out[c] = cur_in[c];
    }
  }
}

template <typename B>
__global__ void __launch_bounds__(512) soft_batch_lsq_tail_kernel(
    const SoftCaseT<B>* __restrict__ cs, int sweeps) {
  const SoftCaseT<B>& C = cs[blockIdx.x];
  const LsqWork& w = C.lw;
  lsq_sweep_tail_body(C.cur, C.e1d, C.e2d, w.cur2, w.ex2, w.ey2, C.tgt, C.z, C.z, 2 * C.nr, C.nz,
                      w.list, w.count, w.counters, sweeps, w.tail, w.cap);
}

template <typename B>
hipError_t launch_soft_batch_lsq_list(const SoftCaseT<B>* cases, const SoftBatchDims& d,
                                      hipStream_t s) {
  soft_batch_lsq_list_kernel<B><<<d.ndbl_blocks, 256, 0, s>>>(cases, d.dbl_blocks);
  return hipGetLastError();
}

template <typename B>
hipError_t launch_soft_batch_lsq_sweep(const SoftCaseT<B>* cases, const SoftBatchDims& d, int sweep,
                                       hipStream_t s) {
  constexpr int kBlocksPerCase = 64;  // a case's band: up to ~16k cells, one per thread
  soft_batch_lsq_sweep_kernel<B><<<dim3(kBlocksPerCase, d.ncase), 256, 0, s>>>(cases, sweep);
  return hipGetLastError();
}

template <typename B>
hipError_t launch_soft_batch_lsq_tail(const SoftCaseT<B>* cases, const SoftBatchDims& d, int sweeps,
                                      hipStream_t s) {
  soft_batch_lsq_tail_kernel<B><<<d.ncase, 512, 0, s>>>(cases, sweeps);
  return hipGetLastError();
}

template hipError_t launch_soft_batch_lsq_list<double>(const SoftCaseT<double>*,
                                                       const SoftBatchDims&, hipStream_t);
template hipError_t launch_soft_batch_lsq_list<float>(const SoftCaseT<float>*,
                                                      const SoftBatchDims&, hipStream_t);
template hipError_t launch_soft_batch_lsq_sweep<double>(const SoftCaseT<double>*,
                                                        const SoftBatchDims&, int, hipStream_t);
template hipError_t launch_soft_batch_lsq_sweep<float>(const SoftCaseT<float>*,
                                                       const SoftBatchDims&, int, hipStream_t);
template hipError_t launch_soft_batch_lsq_tail<double>(const SoftCaseT<double>*,
                                                       const SoftBatchDims&, int, hipStream_t);
template hipError_t launch_soft_batch_lsq_tail<float>(const SoftCaseT<float>*,
                                                      const SoftBatchDims&, int, hipStream_t);

__global__ void fill_u32_kernel(unsigned* p, int n, unsigned v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

hipError_t launch_lsq_sweep(int16_t* cur, const int16_t* tgt, double* ex, double* ey,
                            const double* gx, const double* gy, int n0, int n1, LsqWork w,
                            int sweep, hipStream_t s) {
  dim3 grid(ceil_div(n1, 256), n0);
  lsq_fit_kernel<<<grid, 256, 0, s>>>(cur, tgt, ex, ey, gx, gy, n0, n1, w.nx, w.ny, w.pend,
                                      w.counters + sweep,
                                      sweep > 0 ? w.counters + sweep - 1 : nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int64_t n = (int64_t)n0 * n1;
  lsq_commit_kernel<<<ceil_div(n, 256), 256, 0, s>>>(cur, ex, ey, w.nx, w.ny, w.pend, n,
                                                      w.counters + sweep);
  return hipGetLastError();
}

hipError_t launch_lsq_list(const int16_t* cur, const int16_t* tgt, int64_t n, LsqWork w,
                           hipStream_t s) {
  lsq_list_kernel<<<ceil_div(n, 256), 256, 0, s>>>(cur, tgt, n, w.list, w.count);
  return hipGetLastError();
}

hipError_t launch_lsq_sweep_list(int16_t* cur, const int16_t* tgt, double* ex, double* ey,
                                 const double* gx, const double* gy, int n0, int n1, LsqWork w,
                                 int sweep, hipStream_t s) {
  constexpr int kBlocks = 128;  // the band's cells, strided (a few thousand here)
  lsq_fit_list_kernel<<<kBlocks, 256, 0, s>>>(cur, tgt, ex, ey, gx, gy, n0, n1, w.list, w.count,
                                              w.nx, w.ny, w.pend, w.counters + sweep,
                                              sweep > 0 ? w.counters + sweep - 1 : nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  lsq_commit_list_kernel<<<kBlocks, 256, 0, s>>>(cur, ex, ey, w.nx, w.ny, w.pend, w.list, w.count,
                                                 w.counters + sweep);
  return hipGetLastError();
}

hipError_t launch_lsq_sweep_pingpong(const int16_t* cur_in, const double* ex_in,
                                     const double* ey_in, int16_t* cur_out, double* ex_out,
                                     double* ey_out, const int16_t* tgt, const double* gx,
                                     const double* gy, int n0, int n1, LsqWork w, int sweep,
                                     hipStream_t s) {
  constexpr int kBlocks = 128;
  lsq_sweep_pingpong_kernel<<<kBlocks, 256, 0, s>>>(
      cur_in, ex_in, ey_in, cur_out, ex_out, ey_out, tgt, gx, gy, n0, n1, w.list, w.count,
      w.counters + sweep, sweep > 0 ? w.counters + sweep - 1 : nullptr);
  return hipGetLastError();
}

hipError_t launch_lsq_sweep_tail(int16_t* cur, double* ex, double* ey, const int16_t* tgt,
                                 const double* gx, const double* gy, int n0, int n1, LsqWork w,
                                 int sweeps, hipStream_t s) {
  lsq_sweep_tail_kernel<<<1, 512, 0, s>>>(cur, ex, ey, w.cur2, w.ex2, w.ey2, tgt, gx, gy, n0, n1,
                                           w.list, w.count, w.counters, sweeps, w.tail, w.cap);
  return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned* p, int n, unsigned v, hipStream_t s) {
  fill_u32_kernel<<<ceil_div(n, 256), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

}  // namespace axs

using namespace axs;

extern "C" {

axs_status axs_extrapolate_lsq_first_order_f64(int16_t* current_flag, const int16_t* target_flag,
                                               double* eta_x, double* eta_y,
                                               const double* grid_x, const double* grid_y,
                                               int n0, int n1, int max_sweeps,
                                               int* sweeps_done_host, void* stream) {
  AXS_CHECK_ARG(current_flag && target_flag && eta_x && eta_y && grid_x && grid_y,
                "null pointer");
  AXS_CHECK_ARG(n0 >= 1 && n1 >= 1, "bad shape");
  AXS_CHECK_ARG(max_sweeps >= 1 && max_sweeps <= 1 << 20, "bad max_sweeps");
  hipStream_t s = as_stream(stream);
  const int64_t n = (int64_t)n0 * n1;
  // Sweeps are issued in batches; a sweep whose predecessor found nothing
  // exits at once, so the host only reads the counters once per batch.
  constexpr int kBatch = 8;
  LsqWork w{};
  AXS_HIP(hipMalloc(reinterpret_cast<void**>(&w.nx), sizeof(double) * n));
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&w.ny), sizeof(double) * n);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&w.pend), n);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&w.counters), sizeof(unsigned) * kBatch);
  int done = 0;
  bool finished = false;
  std::vector<unsigned> h(kBatch);
  while (e == hipSuccess && !finished && done < max_sweeps) {
    e = launch_fill_u32(w.counters, kBatch, 0u, s);
    const int nb = std::min(kBatch, max_sweeps - done);
    for (int q = 0; q < nb && e == hipSuccess; ++q)
      e = launch_lsq_sweep(current_flag, target_flag, eta_x, eta_y, grid_x, grid_y, n0, n1, w,
                           q, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(h.data(), w.counters, sizeof(unsigned) * kBatch, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (int q = 0; q < nb && e == hipSuccess; ++q) {
      if (h[q] == 0) {
        finished = true;
        break;
      }
      ++done;
    }
  }
  (void)hipFree(w.nx);
  (void)hipFree(w.ny);
  (void)hipFree(w.pend);
  (void)hipFree(w.counters);
  if (e != hipSuccess) {
    set_error(std::string("lsq extrapolation: ") + hipGetErrorString(e));
    return AXS_ERR_HIP;
  }
  if (sweeps_done_host) *sweeps_done_host = done;
  return AXS_OK;
}

}  // extern "C"
