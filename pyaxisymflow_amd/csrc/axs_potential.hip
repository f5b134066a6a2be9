// Velocity potential of FastDiagonalisationPotentialSolver
// (kernels/FastDiagonalisationPotentialSolver.py:6-145) and the velocity of
// compute_velocity_from_phi (kernels/compute_velocity_from_phi.py:4-17).
//
// The solve restates A_r phi + phi P_z^T = rhs as
//   1. row pass:    rhs_hat = DCT-II_z(rhs)        (the Neumann P_z's eigenbasis)
//   2. column pass: (A_r + lambda_k) phi_hat_k = rhs_hat_k for every z-mode k
//   3. row pass:    phi = DCT-II^{-1}_z(phi_hat)
// with P_z = [1, -2, 1] / dx^2 and Neumann corners (:82-83), eigenvalues
// lambda_k = -(4 / dx^2) sin^2(pi k / (2 nz)), and A_r = P_r + D_r / r with
// P_r[-1, -1] = -1 / dx^2, D_r[-1, -2] = 0 (:84-86), built with the same
// floating-point operations.  Constants are in the null space of both, so
// mode 0's radial system is singular: the reference divides by a round-off
// eigenvalue there (:114-117) and phi carries an arbitrary constant (a huge
// one unless the rhs is compatible).  Here mode 0 is solved in the gauge
// phi_hat_0[nr - 1] = 0: its last equation is replaced by that condition,
// which leaves rows 0 .. nr-2, a nonsingular system (the last one strictly
// diagonally dominant once the coupling to row nr-1 is gone).  For a
// compatible rhs this is an exact solution; u = grad phi does not see the
// constant.  (The transforms and the column pass are the Stokes plan's; this
// file keeps the library's default of no FMA contraction, so the velocity
// stencil below rounds as the reference does.)

#include "axs_potential.hpp"

#include <new>
#include <vector>

namespace axs {

// mode 0 of the spectrum into a contiguous column, its last entry replaced
// by the gauge condition's right-hand side (0)
__global__ void mode0_gather_kernel(const double* __restrict__ spec, double* __restrict__ col,
                                    int nr, int nz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nr) col[i] = (i == nr - 1) ? 0.0 : spec[(int64_t)i * nz];
}

__global__ void mode0_scatter_kernel(double* __restrict__ spec, const double* __restrict__ col,
                                     int nr, int nz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nr) spec[(int64_t)i * nz] = col[i];
}

// compute_velocity_from_phi.py:4-17: u_z = d phi / dz, u_r = d phi / dr,
// central differences inside, second-order one-sided at the edges; the
// reference's operation order and IEEE division
__global__ void velocity_from_phi_kernel(double* __restrict__ u_z, double* __restrict__ u_r,
                                         const double* __restrict__ phi, int nr, int nz,
                                         double dx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  const int64_t c = (int64_t)i * nz + j;
  const double two_dx = 2.0 * dx;
  double uz;
  if (j == 0) {
    const double t = 4.0 * phi[c + 1];
    const double q = 3.0 * phi[c];
    uz = ((-phi[c + 2] + t) - q) / two_dx;
  } else if (j == nz - 1) {
    const double t = 4.0 * phi[c - 1];
    const double q = 3.0 * phi[c];
    uz = ((phi[c - 2] - t) + q) / two_dx;
  } else {
    uz = (phi[c + 1] - phi[c - 1]) / two_dx;
  }
  double ur;
  if (i == 0) {
    const double t = 4.0 * phi[c + nz];
    const double q = 3.0 * phi[c];
    ur = ((-phi[c + 2 * nz] + t) - q) / two_dx;
  } else if (i == nr - 1) {
    const double t = 4.0 * phi[c - nz];
    const double q = 3.0 * phi[c];
    ur = ((phi[c - 2 * nz] - t) + q) / two_dx;
  } else {
    ur = (phi[c + nz] - phi[c - nz]) / two_dx;
  }
  u_z[c] = uz;
  u_r[c] = ur;
}

static hipError_t upload_vec(double** dev, const std::vector<double>& h) {
  hipError_t e = hipMalloc(reinterpret_cast<void**>(dev), sizeof(double) * h.size());
  if (e != hipSuccess) return e;
  return hipMemcpy(*dev, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice);
}

hipError_t PotentialPlan::init(int nr, int nz, double dx) {
  pl.tune_key = 2;
  hipError_t e = pl.init_transforms(nr, nz, dx);
  if (e != hipSuccess) return e;
  const std::vector<double>& r = pl.host_r;
  // :45-86: inv_dx2 * {1, -2, 1} + inv_2dx * {-1, 1} / r, then the Neumann rows
  const double inv_dx2 = 1.0 / dx / dx, inv_2dx = 1.0 / 2.0 / dx;
  std::vector<double> a(nr, 0.0), b(nr), c(nr, 0.0), lam(nz);
  for (int i = 0; i < nr; ++i) {
    b[i] = inv_dx2 * -2.0 + 0.0;
    if (i > 0) a[i] = inv_dx2 * 1.0 + (inv_2dx * -1.0) / r[i];
    if (i < nr - 1) c[i] = inv_dx2 * 1.0 + (inv_2dx * 1.0) / r[i];
  }
  b[nr - 1] = -inv_dx2 + 0.0;      // P_r[-1, -1] = -inv_dx2, D_r[-1, -1] = 0
  a[nr - 1] = inv_dx2 * 1.0 + 0.0;  // D_r[-1, -2] = 0
  for (int k = 0; k < nz; ++k) {
    const double sn = std::sin(M_PI * (double)k / (2.0 * nz));
    lam[k] = -(4.0 / (dx * dx)) * sn * sn;
  }
  if ((e = pl.set_tridiag(a, b, c, lam)) != hipSuccess) return e;
  std::vector<double> ga = a, gb = b, gc = c, gl(1, 0.0);
  ga[nr - 1] = 0.0;
  gb[nr - 1] = 1.0;
  if ((e = upload_vec(&a0, ga)) != hipSuccess) return e;
  if ((e = upload_vec(&b0, gb)) != hipSuccess) return e;
  if ((e = upload_vec(&c0, gc)) != hipSuccess) return e;
  if ((e = upload_vec(&lam0, gl)) != hipSuccess) return e;
  if ((e = upload_vec(&cseg0, pivot_table(ga, gb, gc, gl, nr, 1, 16))) != hipSuccess) return e;
  return hipMalloc(reinterpret_cast<void**>(&col0), sizeof(double) * nr);
}

hipError_t launch_velocity_from_phi(double* u_z, double* u_r, const double* phi, int nr, int nz,
                                    double dx, hipStream_t s) {
  velocity_from_phi_kernel<<<dim3(ceil_div(nz, 256), nr), 256, 0, s>>>(u_z, u_r, phi, nr, nz, dx);
  return hipGetLastError();
}

PotentialPlan::~PotentialPlan() {
  for (double* p : {a0, b0, c0, lam0, cseg0, col0})
    if (p) (void)hipFree(p);
}

hipError_t PotentialPlan::solve(double* phi, const double* rhs, hipStream_t s) {
  const int nr = pl.nr, nz = pl.nz;
  hipError_t e = launch_dct2_rows<double>(pl, rhs, pl.spec, nullptr, nr, s);
  if (e != hipSuccess) return e;
  mode0_gather_kernel<<<ceil_div(nr, 256), 256, 0, s>>>(pl.spec, col0, nr, nz);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = launch_tridiag_one(col0, a0, b0, c0, lam0, cseg0, nr, s)) != hipSuccess) return e;
  // every mode; mode 0's singular solve is then overwritten by the gauge one
  if ((e = launch_tridiag<double>(pl, pl.spec, s)) != hipSuccess) return e;
  mode0_scatter_kernel<<<ceil_div(nr, 256), 256, 0, s>>>(pl.spec, col0, nr, nz);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return launch_dct3_rows<double>(pl, pl.spec, phi, nr, s);
}

}  // namespace axs

using namespace axs;

struct axs_potential_plan {
  PotentialPlan impl;
};

extern "C" {

axs_status axs_potential_plan_create(int nr, int nz, double dx, int bc_type,
                                     axs_potential_plan** plan_out) {
  AXS_CHECK_ARG(plan_out, "null plan_out");
  *plan_out = nullptr;
  AXS_CHECK_ARG(bc_type == AXS_BC_NEUMANN_Z_AND_R,
                "only homogenous_neumann_along_z_and_r is supported");
  AXS_CHECK_ARG(nr >= 3 && nz >= 4, "need nr >= 3 and nz >= 4");
  AXS_CHECK_ARG(nr <= 4096, "nr > 4096: too tall for the column pass");
  AXS_CHECK_ARG(dx > 0.0, "dx must be positive");
  if (!stokes_nz_supported(nz)) {
    set_error("nz too large for the row transforms (nz <= 8192 with nz/2 2,3,5-smooth, "
              "else nz <= 4096)");
    return AXS_ERR_UNSUPPORTED;
  }
  axs_potential_plan* p = new (std::nothrow) axs_potential_plan();
  if (!p) {
    set_error("out of host memory");
    return AXS_ERR_ALLOC;
  }
  hipError_t e = p->impl.init(nr, nz, dx);
  if (e != hipSuccess) {
    set_error(std::string("potential plan: ") + hipGetErrorString(e));
    delete p;
    return AXS_ERR_HIP;
  }
  *plan_out = p;
  return AXS_OK;
}

axs_status axs_potential_plan_destroy(axs_potential_plan* plan) {
  delete plan;
  return AXS_OK;
}

axs_status axs_potential_solve_f64(axs_potential_plan* plan, double* phi, const double* rhs,
                                   void* stream) {
  AXS_CHECK_ARG(plan && phi && rhs, "null pointer");
  AXS_HIP(plan->impl.solve(phi, rhs, as_stream(stream)));
  return AXS_OK;
}

axs_status axs_velocity_from_phi_f64(double* u_z, double* u_r, const double* phi, int nr, int nz,
                                     double dx, void* stream) {
  AXS_CHECK_ARG(u_z && u_r && phi, "null pointer");
  AXS_CHECK_ARG(nr >= 3 && nz >= 3, "need nr >= 3 and nz >= 3");
  velocity_from_phi_kernel<<<dim3(ceil_div(nz, 256), nr), 256, 0, as_stream(stream)>>>(
      u_z, u_r, phi, nr, nz, dx);
  AXS_LAUNCH_CHECK("velocity_from_phi");
  return AXS_OK;
}

}  // extern "C"
