// Particle remeshing advection with the M'4 kernel (kernels/advect_particle.py:5-68)
// and the 2D particle <-> mesh interpolations it rests on
// (core/src/particles_to_mesh.hpp:71-91, 165-185; mesh_to_particles.hpp:111-131,
// 209-229; particle_kernels/MP4.hpp:14-57).
//
// P2M is a scatter in the reference: particle p adds (w_y w_x) q_p to the
// 4 x 4 cells around it, particles in row-major order.  Here it is a gather
// without atomics: each cell sums the contributions of the particles whose
// stencil covers it, visited in increasing particle index, which is the
// order the reference's sequential loop adds them to that cell, so the sums
// are the reference's bit for bit.  The particle array has the mesh's shape
// (the remeshing step's particles start on the cell centres); a first pass
// measures D = max over particles of |h - home| (h the particle's nearest
// mesh index, home its array index), and a cell then gathers from the
// particles within D + 2 of it.  M2P is a gather per particle in the
// reference already.
//
// Arithmetic and its order follow the headers (no FMA contraction: the
// library's default).
#include "axs_common.hpp"

#include <algorithm>
#include <cmath>

namespace axs {

constexpr int kP2MTR = 16, kP2MTC = 64;  // output tile of one P2M workgroup
constexpr int kP2MMaxDisp = 4;           // largest D the gather takes

// MP4.hpp:26-28 (|x| <= 1) and :35-38 (1 < |x| <= 2)
__device__ __forceinline__ double mp4_f0(double t) { return 1.0 + t * t * (-2.5 + 1.5 * t); }
__device__ __forceinline__ double mp4_f1(double t) {
  return 2.0 + t * (-4.0 + t * (2.5 - 0.5 * t));
}
// weight of stencil point s in {-2, -1, 0, 1} (KernelWrapper.hpp:11-92:
// branches 1, 0, 0, 1) for the scaled position X and nearest index h
__device__ __forceinline__ double mp4_weight(double X, int h, int s) {
  const double t = fabs(X - (((double)h + 0.5) + (double)s));
  return (s == -2 || s == 1) ? mp4_f1(t) : mp4_f0(t);
}
// nearest mesh index of a scaled position (particles_to_mesh_2D.hpp:61-75)
__device__ __forceinline__ int mp4_nearest(double X) {
  const double least = floor(X);
  return (int)(X >= least + 0.5) + (int)least;
}

// D = max |h - home| over the particles; any non-finite position or an
// index beyond int range sets the flag instead.
__global__ void p2m_disp_kernel(const double* __restrict__ px, const double* __restrict__ py,
                                int n0, int n1, double dx, double dy, int* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int d = 0;
  if (k < (int64_t)n0 * n1) {
    const int i = (int)(k % n1), j = (int)(k / n1);
    const double X = px[k] / dx, Y = py[k] / dy;
    if (!(fabs(X) < 1e9 && fabs(Y) < 1e9)) {
      d = 1 << 30;
    } else {
      d = max(abs(mp4_nearest(X) - i), abs(mp4_nearest(Y) - j));
    }
  }
  for (int off = 32; off > 0; off >>= 1) d = max(d, __shfl_xor(d, off));
  if ((threadIdx.x & 63) == 0) atomicMax(out, d);
}

// Two-segment walk of a window [lo, hi] of virtual indices in increasing
// order of the actual index v mod n (periodic) or over [max(lo, 0),
// min(hi, n - 1)] (unbounded); at(q) is the q-th virtual index of the walk.
struct WindowOrder {
  int lo, hi, split;  // virtual order: split..hi, then lo..split-1
  __device__ WindowOrder(int lo_, int hi_, int n, bool periodic) {
    if (!periodic) {
      lo = max(lo_, 0);
      hi = min(hi_, n - 1);
      split = lo;
    } else {
      lo = lo_;
      hi = hi_;
      split = lo_ < 0 ? 0 : (hi_ >= n ? n : lo_);
    }
  }
  __device__ int count() const { return hi >= lo ? hi - lo + 1 : 0; }
  // q-th virtual index in increasing actual order
  __device__ int at(int q) const {
    const int first = hi - split + 1;
    return q < first ? split + q : lo + (q - first);
  }
};

// P2M gather.  Workgroup: a kP2MTR x kP2MTC tile of cells; the particles of
// the window (tile + D + 2 around it, in virtual coordinates, wrapped when
// periodic) are staged in LDS as (X, Y, q, h_x, h_y).
template <bool PERIODIC>
__global__ void __launch_bounds__(256) p2m_gather_kernel(
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ q,
    double* __restrict__ out, int n0, int n1, double dx, double dy,
    const int* __restrict__ disp) {
  extern __shared__ unsigned char smem_raw[];
  const int D = *disp;
  const int wr = kP2MTR + 2 * D + 3, wc = kP2MTC + 2 * D + 3;
  double* sX = reinterpret_cast<double*>(smem_raw);
  double* sY = sX + wr * wc;
  double* sQ = sY + wr * wc;
  int* sHx = reinterpret_cast<int*>(sQ + wr * wc);
  int* sHy = sHx + wr * wc;
  const int I0 = blockIdx.y * kP2MTR, J0 = blockIdx.x * kP2MTC;
  const int r0 = I0 - 1 - D, c0 = J0 - 1 - D;  // virtual window origin
  for (int e = threadIdx.x; e < wr * wc; e += blockDim.x) {
    const int vr = r0 + e / wc, vc = c0 + e % wc;
    int r = vr, c = vc;
    bool ok = true;
    if (PERIODIC) {
      r = ((vr % n0) + n0) % n0;
      c = ((vc % n1) + n1) % n1;
    } else {
      ok = vr >= 0 && vr < n0 && vc >= 0 && vc < n1;
    }
    double X = 0.0, Y = 0.0, Q = 0.0;
    int hx = 0, hy = 0;
    if (ok) {
      const int64_t k = (int64_t)r * n1 + c;
      X = px[k] / dx;
      Y = py[k] / dy;
      Q = q[k];
      hx = mp4_nearest(X);
      hy = mp4_nearest(Y);
    }
    sX[e] = X;
    sY[e] = Y;
    sQ[e] = Q;
    sHx[e] = hx;
    sHy[e] = hy;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kP2MTR * kP2MTC; e += blockDim.x) {
    const int I = I0 + e / kP2MTC, J = J0 + e % kP2MTC;
    if (I >= n0 || J >= n1) continue;
    const WindowOrder rows(I - 1 - D, I + 2 + D, n0, PERIODIC);
    const WindowOrder cols(J - 1 - D, J + 2 + D, n1, PERIODIC);
    double acc = 0.0;
    const int nrow = rows.count(), ncol = cols.count();
    for (int a = 0; a < nrow; ++a) {
      const int lr = rows.at(a) - r0;
      for (int b = 0; b < ncol; ++b) {
        const int e2 = lr * wc + (cols.at(b) - c0);
        int sy = I - sHy[e2], sx = J - sHx[e2];
        if (PERIODIC) {  // the stencil offset that lands on (I, J) after wrapping
          if (sy > n0 / 2) sy -= n0;
          if (sy < -n0 / 2) sy += n0;
          if (sx > n1 / 2) sx -= n1;
          if (sx < -n1 / 2) sx += n1;
        }
        if (sy < -2 || sy > 1 || sx < -2 || sx > 1) continue;
        const double wy = mp4_weight(sY[e2], sHy[e2], sy);
        const double wx = mp4_weight(sX[e2], sHx[e2], sx);
        acc = acc + (wy * wx) * sQ[e2];
      }
    }
    out[(int64_t)I * n1 + J] = acc;
  }
}

// M2P per particle (mesh_to_particles_2D.hpp: rows of x partial sums, then
// the y sum over the four rows; unbounded drops cells outside the mesh).
template <bool PERIODIC>
__global__ void m2p_kernel(const double* __restrict__ fx, const double* __restrict__ fy,
                           const double* __restrict__ px, const double* __restrict__ py,
                           double* __restrict__ ox, double* __restrict__ oy, int m0, int m1,
                           int64_t np, double dx, double dy) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const double X = px[k] / dx, Y = py[k] / dy;
  if (!(fabs(X) < 1e9 && fabs(Y) < 1e9)) {  // no mesh index (the reference's is undefined)
    ox[k] = NAN;
    oy[k] = NAN;
    return;
  }
  const int hx = mp4_nearest(X), hy = mp4_nearest(Y);
  double wx[4], wy[4], ps[4][2];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    wx[s] = mp4_weight(X, hx, s - 2);
    wy[s] = mp4_weight(Y, hy, s - 2);
    ps[s][0] = 0.0;
    ps[s][1] = 0.0;
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    int row = hy + a - 2;
    if (PERIODIC)
      row = ((row % m0) + m0) % m0;  // (row + m0) % m0 wherever that is in range
    else if (row < 0 || row >= m0)
      continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      int col = hx + b - 2;
      if (PERIODIC)
        col = ((col % m1) + m1) % m1;
      else if (col < 0 || col >= m1)
        continue;
      const int64_t m = (int64_t)row * m1 + col;
      ps[a][0] += wx[b] * fx[m];
      ps[a][1] += wx[b] * fy[m];
    }
  }
  double ax = 0.0, ay = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    ax += wy[a] * ps[a][0];
    ay += wy[a] * ps[a][1];
  }
  ox[k] = ax;
  oy[k] = ay;
}

// advect_particle.py:20-27: push the particles of the doubled grid (2 nr, nz)
// by the mirrored velocity (u_z even, u_r odd about the axis) and carry the
// mirrored vorticity (odd).
__global__ void particles_push_kernel(double* __restrict__ zp, double* __restrict__ rp,
                                      double* __restrict__ wp, const double* __restrict__ w,
                                      const double* __restrict__ uz,
                                      const double* __restrict__ ur, int nr, int nz, double dt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;  // row of the doubled grid
  if (j >= nz) return;
  const int64_t d = (int64_t)k * nz + j;
  if (k >= nr) {
    const int64_t c = (int64_t)(k - nr) * nz + j;
    zp[d] = zp[d] + uz[c] * dt;
    rp[d] = rp[d] + ur[c] * dt;
    wp[d] = w[c];
  } else {
    const int64_t c = (int64_t)(nr - 1 - k) * nz + j;
    zp[d] = zp[d] + uz[c] * dt;
    rp[d] = rp[d] + (-ur[c]) * dt;
    wp[d] = -w[c];
  }
}

// :29-35: particles back on the grid, the upper half of the remeshed field
// into the vorticity.
__global__ void particles_reset_kernel(double* __restrict__ zp, double* __restrict__ rp,
                                       const double* __restrict__ Zd,
                                       const double* __restrict__ Rd, double* __restrict__ w,
                                       const double* __restrict__ wd, int nr, int nz) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (j >= nz) return;
  const int64_t d = (int64_t)k * nz + j;
  zp[d] = Zd[d];
  rp[d] = Rd[d];
  if (k >= nr) w[(int64_t)(k - nr) * nz + j] = wd[d];
}

// D into *disp_host (synchronizes s); then the gather.
static hipError_t p2m(const double* px, const double* py, const double* q, double* out, int n0,
                      int n1, double dx, double dy, bool periodic, int* d_disp, int* disp_host,
                      hipStream_t s) {
  hipError_t e = hipMemsetAsync(d_disp, 0, sizeof(int), s);
  if (e != hipSuccess) return e;
  const int64_t n = (int64_t)n0 * n1;
  p2m_disp_kernel<<<ceil_div(n, 256), 256, 0, s>>>(px, py, n0, n1, dx, dy, d_disp);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(disp_host, d_disp, sizeof(int), hipMemcpyDeviceToHost, s)) !=
      hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  const int D = *disp_host;
  // periodic: the window of 2 D + 4 particles per axis must not wrap onto
  // itself, and the stencil offset -2 must stay distinct from +2 (n >= 5)
  const int nmin = std::max(2 * D + 4, 8);
  if (D > kP2MMaxDisp || (periodic && (nmin > n0 || nmin > n1))) return hipErrorInvalidValue;
  const int wr = kP2MTR + 2 * D + 3, wc = kP2MTC + 2 * D + 3;
  const size_t shm = (size_t)wr * wc * (3 * sizeof(double) + 2 * sizeof(int));
  const dim3 grid(ceil_div(n1, kP2MTC), ceil_div(n0, kP2MTR));
  if (periodic)
    p2m_gather_kernel<true><<<grid, 256, shm, s>>>(px, py, q, out, n0, n1, dx, dy, d_disp);
  else
    p2m_gather_kernel<false><<<grid, 256, shm, s>>>(px, py, q, out, n0, n1, dx, dy, d_disp);
  return hipGetLastError();
}

}  // namespace axs

using namespace axs;

extern "C" {

axs_status axs_particles_to_mesh_2d_mp4_f64(const double* px, const double* py, const double* q,
                                            double* out, int n0, int n1, double dx, double dy,
                                            int periodic, int* max_displacement_host,
                                            void* stream) {
  AXS_CHECK_ARG(px && py && q && out, "null pointer");
  AXS_CHECK_ARG(n0 >= 1 && n1 >= 1, "empty grid");
  AXS_CHECK_ARG(dx > 0.0 && dy > 0.0, "dx and dy must be positive");
  int* d_disp = nullptr;
  AXS_HIP(hipMalloc(reinterpret_cast<void**>(&d_disp), sizeof(int)));
  int D = 0;
  const hipError_t e =
      p2m(px, py, q, out, n0, n1, dx, dy, periodic != 0, d_disp, &D, as_stream(stream));
  const hipError_t es = (e == hipSuccess) ? hipStreamSynchronize(as_stream(stream)) : e;
  (void)hipFree(d_disp);
  if (max_displacement_host) *max_displacement_host = D;
  if (e == hipErrorInvalidValue) {
    set_error("particles moved more than " + std::to_string(kP2MMaxDisp) +
              " cells from their home cells (or the periodic mesh is too small for the "
              "stencil): " + std::to_string(D));
    return AXS_ERR_UNSUPPORTED;
  }
  AXS_HIP(es);
  return AXS_OK;
}

axs_status axs_mesh_to_particles_2d_mp4_f64(const double* fx, const double* fy, const double* px,
                                            const double* py, double* ox, double* oy, int m0,
                                            int m1, int64_t np, double dx, double dy,
                                            int periodic, void* stream) {
  AXS_CHECK_ARG(fx && fy && px && py && ox && oy, "null pointer");
  AXS_CHECK_ARG(m0 >= 1 && m1 >= 1 && np >= 0, "bad sizes");
  AXS_CHECK_ARG(dx > 0.0 && dy > 0.0, "dx and dy must be positive");
  if (np == 0) return AXS_OK;
  if (periodic)
    m2p_kernel<true><<<ceil_div(np, 256), 256, 0, as_stream(stream)>>>(fx, fy, px, py, ox, oy, m0,
                                                                        m1, np, dx, dy);
  else
    m2p_kernel<false><<<ceil_div(np, 256), 256, 0, as_stream(stream)>>>(fx, fy, px, py, ox, oy,
                                                                         m0, m1, np, dx, dy);
  AXS_LAUNCH_CHECK("m2p");
  return AXS_OK;
}

axs_status axs_advect_vorticity_via_particles_f64(double* z_particles, double* r_particles,
                                                  double* vort_particles, double* vorticity,
                                                  const double* Z_double, const double* R_double,
                                                  const double* u_z, const double* u_r,
                                                  double* vort_double, int nr, int nz, double dx,
                                                  double dt, int periodic, void* stream) {
  AXS_CHECK_ARG(z_particles && r_particles && vort_particles && vorticity && Z_double &&
                    R_double && u_z && u_r && vort_double,
                "null pointer");
  AXS_CHECK_ARG(nr >= 1 && nz >= 1, "empty grid");
  AXS_CHECK_ARG(dx > 0.0, "dx must be positive");
  hipStream_t s = as_stream(stream);
  const dim3 grid(ceil_div(nz, 256), 2 * nr);
  particles_push_kernel<<<grid, 256, 0, s>>>(z_particles, r_particles, vort_particles, vorticity,
                                             u_z, u_r, nr, nz, dt);
  AXS_LAUNCH_CHECK("particles_push");
  int* d_disp = nullptr;
  AXS_HIP(hipMalloc(reinterpret_cast<void**>(&d_disp), sizeof(int)));
  int D = 0;
  hipError_t e = p2m(z_particles, r_particles, vort_particles, vort_double, 2 * nr, nz, dx, dx,
                     periodic != 0, d_disp, &D, s);
  if (e == hipSuccess) {
    particles_reset_kernel<<<grid, 256, 0, s>>>(z_particles, r_particles, Z_double, R_double,
                                                vorticity, vort_double, nr, nz);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d_disp);
  if (e == hipErrorInvalidValue) {
    set_error("particles moved more than " + std::to_string(kP2MMaxDisp) +
              " cells in one step (or the periodic mesh is too small for the stencil): " +
              std::to_string(D));
    return AXS_ERR_UNSUPPORTED;
  }
  AXS_HIP(e);
  return AXS_OK;
}

}  // extern "C"
