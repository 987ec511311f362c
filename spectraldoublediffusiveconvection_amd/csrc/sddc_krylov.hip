// Matrix-free Newton-Krylov steady solve on the device (float64 throughout):
// restarted GMRES(m) with classical Gram-Schmidt applied twice (CGS2) on a
// Krylov basis resident in HBM, around the step and the linearised step.
// Newton follows _Newton (Main.py:430-554): residual F(X) = Step(X) - X,
// Jacobian-vector product J v = PDFX(v) = Step'(X) v - v (Main.py:498-521),
// X <- X - dv, converged when ||dv|| / ||X|| <= tol within 5 iterations.
//
// Reductions are deterministic: each dot product is summed per workgroup in
// a fixed tree, and the per-workgroup partials in a fixed order, so a solve
// gives the same bits on every run.  The small Hessenberg least-squares
// problem (Givens rotations, back substitution) is solved on the host from
// the j+1 projections copied back per Arnoldi step.
#include <cmath>
#include <vector>

#include "sddc_internal.h"

struct sddc_krylov {
  sddc_plan* plan = nullptr;
  int n = 0;           // vector length 3N
  int m = 0;           // restart length
  double* V = nullptr;     // (m+1) x n basis
  double* w = nullptr;     // n
  double* r = nullptr;     // n residual / work
  double* f = nullptr;     // n Newton residual
  double* t = nullptr;     // n step output / work
  double* partial = nullptr;   // (m+2) x kBlocks partial sums
  double* h = nullptr;         // 2(m+2)+1: first and second Gram-Schmidt projections, a norm
  std::vector<double> hh;      // host copy
};

namespace sddc {

constexpr int kRedThreads = 256;
constexpr int kBlocks = 240;   // workgroups of the reductions

// partial[i][blk] = sum over this workgroup's chunk of A_i . x for i < nv,
// where A_i = V + i*n (or y when V is null: partial norm of y).
__global__ __launch_bounds__(kRedThreads) void k_dots(const double* __restrict__ V,
                                                      const double* __restrict__ x, int n,
                                                      int nv, double* __restrict__ partial) {
  __shared__ double red[kRedThreads];
  const int chunk = (n + gridDim.x - 1) / gridDim.x;
  const int e0 = blockIdx.x * chunk;
  const int e1 = min(n, e0 + chunk);
  for (int i = 0; i < nv; ++i) {
    const double* a = V ? V + (size_t)i * n : x;
    double s = 0.0;
    for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) s = fma(a[e], x[e], s);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = kRedThreads / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)i * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

// out[i] = (acc ? out[i] : 0) + sum_blk partial[i][blk], fixed order.
__global__ void k_reduce(const double* __restrict__ partial, int nblk, int nv, int acc,
                         double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partial[(size_t)i * nblk + b];
  out[i] = acc ? out[i] + s : s;
}

// y = alpha*y + beta * sum_{i<nv} c_i V_i   (c on the device)
__global__ void k_combine(double* __restrict__ y, const double* __restrict__ V,
                          const double* __restrict__ c, int n, int nv, double alpha,
                          double beta) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)n;
       e += (size_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int i = 0; i < nv; ++i) s = fma(c[i], V[(size_t)i * n + e], s);
    y[e] = alpha * y[e] + beta * s;
  }
}

// z = a*x + b*y
__global__ void k_axpby(double* __restrict__ z, const double* __restrict__ x,
                        const double* __restrict__ y, int n, double a, double b) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)n;
       e += (size_t)gridDim.x * blockDim.x)
    z[e] = a * x[e] + (y ? b * y[e] : 0.0);
}

static int grid_for(int n) { return std::min(1024, (n + 255) / 256); }

static int axpby(double* z, const double* x, const double* y, int n, double a, double b,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(256), 0, s, z, x, y, n, a, b);
  return hip_check(hipGetLastError(), "k_axpby");
}

// h[0..nv) (device) = V_i . x, optionally accumulated
static int dots(sddc_krylov& k, const double* V, const double* x, int nv, int acc, double* out,
                hipStream_t s) {
  hipLaunchKernelGGL(k_dots, dim3(kBlocks), dim3(kRedThreads), 0, s, V, x, k.n, nv, k.partial);
  if (int rc = hip_check(hipGetLastError(), "k_dots")) return rc;
  hipLaunchKernelGGL(k_reduce, dim3((nv + 63) / 64), dim3(64), 0, s, k.partial, kBlocks, nv, acc,
                     out);
  return hip_check(hipGetLastError(), "k_reduce");
}

static double* proj2(sddc_krylov& k) { return k.h + (k.m + 2); }
static double* norm_slot(sddc_krylov& k) { return k.h + 2 * (k.m + 2); }

static int norm2(sddc_krylov& k, const double* x, double* host, hipStream_t s) {
  if (int rc = dots(k, nullptr, x, 1, 0, norm_slot(k), s)) return rc;
  double v = 0.0;
  SDDC_HIP(hipMemcpyAsync(&v, norm_slot(k), sizeof(double), hipMemcpyDeviceToHost, s));
  SDDC_HIP(hipStreamSynchronize(s));
  *host = std::sqrt(v);
  return SDDC_OK;
}

// Jacobian-vector product out = J v of F(X) = Step(X) - X.
static int jv(sddc_krylov& k, const double* Ra, const double* Ras, const double* X,
              const double* v, double* out, int fd, hipStream_t s) {
  int rc;
  if (!fd) {
    if ((rc = sddc_step_linearized(k.plan, 1, Ra, Ras, v, X, out, s))) return rc;
    return axpby(out, out, v, k.n, 1.0, -1.0, s);
  }
  // finite difference of the step: (Step(X + eps v) - Step(X)) / eps - v,
  // Step(X) is kept in k.t by the caller
  double nv, nx;
  if ((rc = norm2(k, v, &nv, s))) return rc;
  if ((rc = norm2(k, X, &nx, s))) return rc;
  if (nv == 0.0) return axpby(out, v, nullptr, k.n, 0.0, 0.0, s);
  const double eps = std::sqrt(2.2204460492503131e-16) * (1.0 + nx) / nv;
  if ((rc = axpby(k.r, X, v, k.n, 1.0, eps, s))) return rc;
  if ((rc = sddc_step(k.plan, 1, Ra, Ras, k.r, out, 1, 0, s))) return rc;
  if ((rc = axpby(out, out, k.t, k.n, 1.0 / eps, -1.0 / eps, s))) return rc;
  return axpby(out, out, v, k.n, 1.0, -1.0, s);
}

}  // namespace sddc

using namespace sddc;

extern "C" {

int sddc_krylov_create(sddc_plan* plan, int restart, sddc_krylov** out) {
  if (!plan || !out) return fail(SDDC_ERR_INVALID, "null argument");
  if (restart < 1 || restart > 2000) return fail(SDDC_ERR_INVALID, "restart must be in [1, 2000]");
  *out = nullptr;
  sddc_krylov* k = new sddc_krylov();
  k->plan = plan;
  k->n = 3 * plan->p.K * plan->p.nr;
  k->m = restart;
  k->hh.assign(restart + 2, 0.0);
  const size_t n = (size_t)k->n;
  auto A = [&](double** p, size_t count) {
    return hip_check(hipMalloc((void**)p, count * sizeof(double)), "krylov workspace");
  };
  int rc = A(&k->V, (size_t)(restart + 1) * n);
  if (!rc) rc = A(&k->w, n);
  if (!rc) rc = A(&k->r, n);
  if (!rc) rc = A(&k->f, n);
  if (!rc) rc = A(&k->t, n);
  if (!rc) rc = A(&k->partial, (size_t)(restart + 2) * kBlocks);
  if (!rc) rc = A(&k->h, 2 * ((size_t)restart + 2) + 1);
  if (!rc && plan->p.max_batch < 1) rc = sddc_plan_reserve(plan, 1);
  if (rc) {
    sddc_krylov_destroy(k);
    return rc;
  }
  *out = k;
  return SDDC_OK;
}

int sddc_krylov_destroy(sddc_krylov* k) {
  if (!k) return SDDC_OK;
  double* bufs[] = {k->V, k->w, k->r, k->f, k->t, k->partial, k->h};
  for (double* b : bufs)
    if (b) (void)hipFree(b);
  delete k;
  return SDDC_OK;
}

int sddc_gmres(sddc_krylov* k, const double* Ra, const double* Ra_s, const double* X,
               const double* b, double* x, const sddc_krylov_opts* opts, sddc_krylov_info* info,
               void* stream) {
  if (!k || !X || !b || !x || !opts) return fail(SDDC_ERR_INVALID, "null argument");
  hipStream_t s = (hipStream_t)stream;
  const int n = k->n, m = k->m, fd = opts->finite_difference;
  int rc, total = 0;
  double bnorm, rnorm = 0.0;
  if ((rc = norm2(*k, b, &bnorm, s))) return rc;
  const double target = std::max(opts->atol, opts->rtol * bnorm);
  if (fd && (rc = sddc_step(k->plan, 1, Ra, Ra_s, X, k->t, 1, 0, s))) return rc;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  int converged = 0, restarts = 0;
  for (; restarts < std::max(1, opts->max_restarts); ++restarts) {
    // r = b - J x
    if ((rc = jv(*k, Ra, Ra_s, X, x, k->r, fd, s))) return rc;
    if ((rc = axpby(k->r, b, k->r, n, 1.0, -1.0, s))) return rc;
    double beta;
    if ((rc = norm2(*k, k->r, &beta, s))) return rc;
    rnorm = beta;
    if (beta <= target) {
      converged = 1;
      break;
    }
    if ((rc = axpby(k->V, k->r, nullptr, n, 1.0 / beta, 0.0, s))) return rc;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int j = 0;
    for (; j < m; ++j) {
      double* vj = k->V + (size_t)j * n;
      double* vn = k->V + (size_t)(j + 1) * n;
      if ((rc = jv(*k, Ra, Ra_s, X, vj, vn, fd, s))) return rc;
      ++total;
      // CGS2: h = V^T w, w -= V h, twice (projections accumulated)
      if ((rc = dots(*k, k->V, vn, j + 1, 0, k->h, s))) return rc;
      hipLaunchKernelGGL(k_combine, dim3(grid_for(n)), dim3(256), 0, s, vn, k->V, k->h, n, j + 1,
                         1.0, -1.0);
      if ((rc = hip_check(hipGetLastError(), "k_combine"))) return rc;
      if ((rc = dots(*k, k->V, vn, j + 1, 0, proj2(*k), s))) return rc;
      hipLaunchKernelGGL(k_combine, dim3(grid_for(n)), dim3(256), 0, s, vn, k->V, proj2(*k), n,
                         j + 1, 1.0, -1.0);
      if ((rc = hip_check(hipGetLastError(), "k_combine"))) return rc;
      SDDC_HIP(hipMemcpyAsync(k->hh.data(), k->h, (j + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
      std::vector<double> h2(j + 1);
      SDDC_HIP(hipMemcpyAsync(h2.data(), proj2(*k), (j + 1) * sizeof(double),
                              hipMemcpyDeviceToHost, s));
      double hn;
      if ((rc = norm2(*k, vn, &hn, s))) return rc;   // synchronises the copies too
      for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = k->hh[i] + h2[i];
      H[(size_t)(j + 1) * m + j] = hn;
      if (hn > 0.0 && (rc = axpby(vn, vn, nullptr, n, 1.0 / hn, 0.0, s))) return rc;
      // Givens rotations of the new column
      for (int i = 0; i < j; ++i) {
        const double a = H[(size_t)i * m + j], c = H[(size_t)(i + 1) * m + j];
        H[(size_t)i * m + j] = cs[i] * a + sn[i] * c;
        H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * c;
      }
      const double a = H[(size_t)j * m + j], c = H[(size_t)(j + 1) * m + j];
      const double den = std::hypot(a, c);
      cs[j] = den > 0.0 ? a / den : 1.0;
      sn[j] = den > 0.0 ? c / den : 0.0;
      H[(size_t)j * m + j] = den;
      H[(size_t)(j + 1) * m + j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      rnorm = std::fabs(g[j + 1]);
      if (rnorm <= target || hn == 0.0) {
        ++j;
        break;
      }
    }
    if (j > m) j = m;
    // y = H(0:j,0:j)^-1 g(0:j); x += V y
    for (int i = j - 1; i >= 0; --i) {
      double acc = g[i];
      for (int q = i + 1; q < j; ++q) acc -= H[(size_t)i * m + q] * y[q];
      y[i] = acc / H[(size_t)i * m + i];
    }
    SDDC_HIP(hipMemcpyAsync(k->h, y.data(), j * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_combine, dim3(grid_for(n)), dim3(256), 0, s, x, k->V, k->h, n, j, 1.0,
                       1.0);
    if ((rc = hip_check(hipGetLastError(), "k_combine"))) return rc;
    SDDC_HIP(hipStreamSynchronize(s));   // y is a host buffer
    if (rnorm <= target) {
      converged = 1;
      ++restarts;
      break;
    }
  }
  if (info) {
    info->converged = converged;
    info->iterations = total;
    info->restarts = restarts;
    info->residual = rnorm;
    info->rhs_norm = bnorm;
  }
  return SDDC_OK;
}

int sddc_newton(sddc_krylov* k, const double* Ra, const double* Ra_s, double* X,
                const sddc_krylov_opts* opts, sddc_newton_info* info, void* stream) {
  if (!k || !X || !opts) return fail(SDDC_ERR_INVALID, "null argument");
  hipStream_t s = (hipStream_t)stream;
  const int n = k->n;
  const int max_newton = opts->max_newton > 0 ? opts->max_newton : 5;
  int rc, it = 0, krylov_total = 0, gmres_ok = 1;
  double err = 1.0;
  sddc_krylov_opts go = *opts;
  while (err > opts->tol_newton && it < max_newton) {
    // F(X) = Step(X) - X (the step applies the plan's symmetry mask)
    if (k->plan->p.symmetric && (rc = launch_mask(k->plan->p, 1, X, X, s))) return rc;
    if ((rc = sddc_step(k->plan, 1, Ra, Ra_s, X, k->f, 1, 0, s))) return rc;
    if ((rc = axpby(k->f, k->f, X, n, 1.0, -1.0, s))) return rc;
    double fnorm;
    if ((rc = norm2(*k, k->f, &fnorm, s))) return rc;
    go.atol = opts->tol_gmres * fnorm;   // atol = tol_gmres * ||b|| (Main.py:534)
    if ((rc = axpby(k->w, k->f, nullptr, n, 0.0, 0.0, s))) return rc;   // dv = 0
    sddc_krylov_info gi{};
    if ((rc = sddc_gmres(k, Ra, Ra_s, X, k->f, k->w, &go, &gi, s))) return rc;
    krylov_total += gi.iterations;
    gmres_ok = gi.converged;   // the last solve's status decides, as in Main.py:541
    if ((rc = axpby(X, X, k->w, n, 1.0, -1.0, s))) return rc;
    double dn, xn;
    if ((rc = norm2(*k, k->w, &dn, s))) return rc;
    if ((rc = norm2(*k, X, &xn, s))) return rc;
    err = xn > 0.0 ? dn / xn : INFINITY;
    ++it;
  }
  if (info) {
    // Main.py:541-544: failure when all iterations were used or the last
    // Krylov solve did not converge
    info->iterations = it;
    info->converged = (it < max_newton && gmres_ok) ? 1 : 0;
    info->error = err;
    info->krylov_iterations = krylov_total;
  }
  return SDDC_OK;
}

}  // extern "C"
