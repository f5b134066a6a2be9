// Implicit Euler diffusion of ImplicitEulerDiffusionStepper
// (kernels/implicit_diffusion_solver.py:6-143): solves
//   (I - nu dt L) w' = w,   L w = A_r w + w P_z^T,
// P_z = [1, -2, 1] / dx^2 with zero ghost values at both z ends (:46-52),
// A_r = P_r + D_r / r - I / r^2 (:53-70), restated as
//   1. row pass:    what = DST-I_z(w)             (S w^T, S the orthonormal DST-I)
//   2. column pass: (1 - nu dt (A_r + lambda_k)) w'hat_k = what_k per z-mode k
//   3. row pass:    w' = DST-I_z(w'hat)           (S is symmetric and S S = I)
// The Dirichlet P_z has the eigenvectors sin(pi (j+1)(k+1) / (nz+1)) and the
// eigenvalues lambda_k = -(4 / dx^2) sin^2(pi (k+1) / (2 (nz+1))).  The
// radial matrix is tridiagonal and strictly diagonally dominant (diagonal
// 1 + nu dt (2 / dx^2 + 1 / r^2 - lambda_k) against off-diagonals summing to
// nu dt 2 / dx^2), so the column pass of the Stokes solve (segmented Thomas
// elimination without pivoting) applies with other coefficient tables.
// The row passes are fast sine transforms: one complex DFT of nz + 1 points
// per row (axs_anysize.hip; 257 is prime and 2049 = 3 x 683 for the
// power-of-two grids, so those lengths take Bluestein's chirp-z form).
#pragma clang fp contract(fast)

#include "axs_common.hpp"
#include "axs_implicit.hpp"

#include <cmath>
#include <cstdlib>
#include <new>
#include <vector>

namespace axs {

ImplicitPlan::~ImplicitPlan() {
  if (em) (void)hipFree(em);
}

hipError_t ImplicitPlan::init(int nr, int nz, double dx, double nu_dt_, int batch) {
  nu_dt = nu_dt_;
  tri.nr = nr;
  tri.nz = nz;
  tri.dx = dx;
  tri.tune_key = 1;
  // radial coordinate as implicit_diffusion_solver.py:22-26 (numpy.linspace)
  std::vector<double> r(nr);
  const double start = dx / 2, stop = nr * dx - dx / 2;
  const double step = (stop - start) / (nr - 1);
  for (int i = 0; i < nr; ++i) r[i] = i * step + start;
  r[nr - 1] = stop;
  // A_r entries as :44-70 form them: inv_dx2 * {1, -2, 1} plus
  // inv_2dx * {-1, 1} / r on the off-diagonals and -1 / r^2 on the diagonal
  const double inv_dx2 = 1.0 / dx / dx, inv_2dx = 1.0 / 2.0 / dx;
  std::vector<double> a(nr, 0.0), b(nr), c(nr, 0.0), lam(nz);
  for (int i = 0; i < nr; ++i) {
    const double sub = inv_dx2 * 1.0 + (inv_2dx * -1.0) / r[i];
    const double sup = inv_dx2 * 1.0 + (inv_2dx * 1.0) / r[i];
    const double dia = inv_dx2 * -2.0 - 1.0 / (r[i] * r[i]);
    b[i] = 1.0 - nu_dt * dia;
    if (i > 0) a[i] = -nu_dt * sub;
    if (i < nr - 1) c[i] = -nu_dt * sup;
  }
  for (int k = 0; k < nz; ++k) {
    const double sn = std::sin(M_PI * (double)(k + 1) / (2.0 * (nz + 1)));
    lam[k] = -nu_dt * (-(4.0 / (dx * dx)) * sn * sn);
  }
  hipError_t e = tri.set_tridiag(a, b, c, lam);
  if (e == hipSuccess && batch > 1) {  // the spectral buffer of a whole batch
    (void)hipFree(tri.spec);
    tri.spec = nullptr;
    e = hipMalloc(reinterpret_cast<void**>(&tri.spec), sizeof(double) * (size_t)nr * nz * batch);
  }
  // DST-I through a DFT of nz + 1 points; em[k] = e^{-i pi k / (nz + 1)}
  if (e == hipSuccess) e = dst.init(nz + 1);
  if (e != hipSuccess) return e;
  std::vector<Cplx<double>> h_em(nz + 1);
  for (int k = 0; k <= nz; ++k) {
    const double ang = -M_PI * (double)k / (double)(nz + 1);
    h_em[k] = {std::cos(ang), std::sin(ang)};
  }
  e = hipMalloc(reinterpret_cast<void**>(&em), sizeof(Cplx<double>) * h_em.size());
  if (e == hipSuccess)
    e = hipMemcpy(em, h_em.data(), sizeof(Cplx<double>) * h_em.size(), hipMemcpyHostToDevice);
  return e;
}

hipError_t ImplicitPlan::launch(double* omega, int batch, hipStream_t s) {
  const int rows = tri.nr * batch;
  hipError_t e = launch_any_dst1_rows(dst, em, omega, tri.spec, rows, tri.nz, s);
  if (e == hipSuccess) e = launch_tridiag<double>(tri, tri.spec, s, batch);
  if (e == hipSuccess) e = launch_any_dst1_rows(dst, em, tri.spec, omega, rows, tri.nz, s);
  return e;
}

}  // namespace axs

using namespace axs;

struct axs_implicit_plan {
  ImplicitPlan impl;
};

extern "C" {

axs_status axs_implicit_diffusion_plan_create(int nr, int nz, double dx, double nu_dt,
                                              axs_implicit_plan** plan_out) {
  AXS_CHECK_ARG(plan_out, "null plan_out");
  *plan_out = nullptr;
  AXS_CHECK_ARG(nr >= 3 && nz >= 2, "need nr >= 3 and nz >= 2");
  AXS_CHECK_ARG(nr <= 4096, "nr > 4096: too tall for the column pass");
  {
    bool blue = false;
    AXS_CHECK_ARG(any_dft_length(nz + 1, &blue) <= kAnyMaxL,
                  "nz too large for the sine transform rows (nz + 1 <= 4096, or <= 8192 when "
                  "nz + 1 is 2,3,5-smooth)");
  }
  AXS_CHECK_ARG(dx > 0.0, "dx must be positive");
  AXS_CHECK_ARG(nu_dt >= 0.0 && std::isfinite(nu_dt), "nu * dt must be finite and >= 0");
  axs_implicit_plan* p = new (std::nothrow) axs_implicit_plan();
  if (!p) {
    set_error("out of host memory");
    return AXS_ERR_ALLOC;
  }
  hipError_t e = p->impl.init(nr, nz, dx, nu_dt);
  if (e != hipSuccess) {
    set_error(std::string("implicit diffusion plan: ") + hipGetErrorString(e));
    delete p;
    return AXS_ERR_HIP;
  }
  *plan_out = p;
  return AXS_OK;
}

axs_status axs_implicit_diffusion_plan_destroy(axs_implicit_plan* plan) {
  delete plan;
  return AXS_OK;
}

axs_status axs_implicit_diffusion_step_f64(axs_implicit_plan* plan, double* omega,
                                           void* stream) {
  AXS_CHECK_ARG(plan && omega, "null pointer");
  AXS_HIP(plan->impl.launch(omega, 1, as_stream(stream)));
  return AXS_OK;
}

}  // extern "C"
