// Matrix-free Newton-Krylov steady solve on the device (float64 throughout):
// restarted GMRES(m) with classical Gram-Schmidt applied twice (CGS2) on a
// Krylov basis resident in HBM, around the step and the linearised step.
// Newton follows _Newton (Main.py:430-554): residual F(X) = Step(X) - X,
// Jacobian-vector product J v = PDFX(v) = Step'(X) v - v (Main.py:498-521),
// X <- X - dv, converged when ||dv|| / ||X|| <= tol within 5 iterations.
//
// The pseudo-arclength continuation step (_ContinC, Main.py:754-950) runs
// the same GMRES on the bordered system of (X, Ra): the Jacobian row block
// [PDFX | dF/dRa] and the arclength row [delta X_dot, (1 - delta) Ra_dot].
//
// Reductions are deterministic: each dot product is summed per workgroup in
// a fixed tree, and the per-workgroup partials in a fixed order, so a solve
// gives the same bits on every run.  The small Hessenberg least-squares
// problem (Givens rotations, back substitution) is solved on the host from
// the j+1 projections copied back per Arnoldi step into pinned host memory
// (no allocation, free or device-wide synchronisation after creation).
#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

#include "sddc_internal.h"

struct sddc_krylov {
  sddc_plan* plan = nullptr;
  int n = 0;           // vector length 3N
  int m = 0;           // restart length
  double* V = nullptr;     // (m+1) x (n+1) basis; row stride = length of the system solved
  double* w = nullptr;     // n+1
  double* r = nullptr;     // n+1 residual / work
  double* f = nullptr;     // n+1 Newton residual / bordered right-hand side
  double* t = nullptr;     // n+1 step output / work
  double* c[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};   // n+1 each, continuation
  double* scal = nullptr;      // device scalars: 1, 0, Ra of the linearisation, a dot product
  double* partial = nullptr;   // (m+2) x kBlocks partial sums
  double* h = nullptr;         // 2(m+2)+1: first and second Gram-Schmidt projections, a norm
  // pinned host mirror (hipHostMalloc, 2(m+2)+4 doubles): projections, then
  // the scalar slot of norm2/get_scalar; every copy is an async DMA on the
  // caller's stream followed by a wait on that stream only
  double* hpin = nullptr;
};

namespace sddc {

constexpr int kRedThreads = 256;
constexpr int kBlocks = 240;   // workgroups of the reductions

// partial[i][blk] = sum over this workgroup's chunk of A_i . x for i < nv,
// where A_i = V + i*ld (or x itself when V is null: partial norm of x).
__global__ __launch_bounds__(kRedThreads) void k_dots(const double* __restrict__ V,
                                                      const double* __restrict__ x, int n,
                                                      int ld, int nv,
                                                      double* __restrict__ partial) {
  __shared__ double red[kRedThreads];
  const int chunk = (n + gridDim.x - 1) / gridDim.x;
  const int e0 = blockIdx.x * chunk;
  const int e1 = min(n, e0 + chunk);
  for (int i = 0; i < nv; ++i) {
    const double* a = V ? V + (size_t)i * ld : x;
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

// y = alpha*y + beta * sum_{i<nv} c_i V_i   (c on the device, V_i = V + i*ld)
__global__ void k_combine(double* __restrict__ y, const double* __restrict__ V,
                          const double* __restrict__ c, int n, int ld, int nv, double alpha,
                          double beta) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)n;
       e += (size_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int i = 0; i < nv; ++i) s = fma(c[i], V[(size_t)i * ld + e], s);
    y[e] = alpha * y[e] + beta * s;
  }
}

// z = a*x + b*y; z may be x or y (in-place updates), so none is restrict
__global__ void k_axpby(double* z, const double* x, const double* y, int n, double a, double b) {
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

// Row stride of the Krylov basis for systems of length len: a multiple of 8
// doubles, so that every basis vector handed to the step kernels is 64-byte
// aligned like the caller's states.
static int basis_ld(int len) { return (len + 7) & ~7; }

// out[0..nv) (device) = V_i . x over len entries (V_i = V + i*basis_ld(len)),
// optionally accumulated
static int dots(sddc_krylov& k, const double* V, const double* x, int len, int nv, int acc,
                double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_dots, dim3(kBlocks), dim3(kRedThreads), 0, s, V, x, len, basis_ld(len), nv,
                     k.partial);
  if (int rc = hip_check(hipGetLastError(), "k_dots")) return rc;
  hipLaunchKernelGGL(k_reduce, dim3((nv + 63) / 64), dim3(64), 0, s, k.partial, kBlocks, nv, acc,
                     out);
  return hip_check(hipGetLastError(), "k_reduce");
}

static double* proj2(sddc_krylov& k) { return k.h + (k.m + 2); }
static double* norm_slot(sddc_krylov& k) { return k.h + 2 * (k.m + 2); }
static double* host_scalar(sddc_krylov& k) { return k.hpin + 2 * (k.m + 2); }

static int norm2(sddc_krylov& k, const double* x, int len, double* host, hipStream_t s) {
  if (int rc = dots(k, nullptr, x, len, 1, 0, norm_slot(k), s)) return rc;
  SDDC_HIP(hipMemcpyAsync(host_scalar(k), norm_slot(k), sizeof(double), hipMemcpyDeviceToHost, s));
  SDDC_HIP(hipStreamSynchronize(s));
  *host = std::sqrt(*host_scalar(k));
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
  if ((rc = norm2(k, v, k.n, &nv, s))) return rc;
  if ((rc = norm2(k, X, k.n, &nx, s))) return rc;
  if (nv == 0.0) return axpby(out, v, nullptr, k.n, 0.0, 0.0, s);
  const double eps = std::sqrt(2.2204460492503131e-16) * (1.0 + nx) / nv;
  if ((rc = axpby(k.r, X, v, k.n, 1.0, eps, s))) return rc;
  if ((rc = sddc_step(k.plan, 1, Ra, Ras, k.r, out, 1, 0, s))) return rc;
  if ((rc = axpby(out, out, k.t, k.n, 1.0 / eps, -1.0 / eps, s))) return rc;
  return axpby(out, out, v, k.n, 1.0, -1.0, s);
}

// out[e] += v[n] * dfmu[e] for e < n; out[n] = delta * (X_dot . v)[0:n] + (1 - delta) mu_dot v[n]
// (the bordered rows of DG, Main.py:912-913).  dot points at the device dot product.
__global__ void k_border(double* __restrict__ out, const double* __restrict__ v,
                         const double* __restrict__ dfmu, const double* __restrict__ dot, int n,
                         double delta, double mudot) {
  const double vmu = v[n];
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)n;
       e += (size_t)gridDim.x * blockDim.x)
    out[e] = fma(vmu, dfmu[e], out[e]);
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = delta * dot[0] + (1.0 - delta) * mudot * vmu;
}

// x[i] = value (a device scalar written in stream order)
__global__ void k_set(double* __restrict__ x, double value) { x[0] = value; }

static int set_scalar(double* x, double value, hipStream_t s) {
  hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s, x, value);
  return hip_check(hipGetLastError(), "k_set");
}

static int get_scalar(sddc_krylov& k, const double* x, double* host, hipStream_t s) {
  SDDC_HIP(hipMemcpyAsync(host_scalar(k), x, sizeof(double), hipMemcpyDeviceToHost, s));
  SDDC_HIP(hipStreamSynchronize(s));
  *host = *host_scalar(k);
  return SDDC_OK;
}

// A linear operator out = A v on vectors of length len.
using ApplyFn = std::function<int(const double* v, double* out)>;

// Restarted GMRES(m) on A x = b (x holds the initial guess): CGS2 Arnoldi
// on the basis V (row stride len), Givens rotations and the triangular solve
// of the small least-squares problem on the host.  Stops at
// ||b - A x|| <= max(atol, rtol ||b||).
static int gmres_core(sddc_krylov& k, int len, const ApplyFn& A, const double* b, double* x,
                      const sddc_krylov_opts& opts, sddc_krylov_info* info, hipStream_t s) {
  const int m = k.m, ld = basis_ld(len);
  int rc, total = 0;
  double bnorm, rnorm = 0.0;
  if ((rc = norm2(k, b, len, &bnorm, s))) return rc;
  const double target = std::max(opts.atol, opts.rtol * bnorm);
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  int converged = 0, restarts = 0;
  for (; restarts < std::max(1, opts.max_restarts); ++restarts) {
    // r = b - A x
    if ((rc = A(x, k.r))) return rc;
    if ((rc = axpby(k.r, b, k.r, len, 1.0, -1.0, s))) return rc;
    double beta;
    if ((rc = norm2(k, k.r, len, &beta, s))) return rc;
    rnorm = beta;
    if (beta <= target) {
      converged = 1;
      break;
    }
    if ((rc = axpby(k.V, k.r, nullptr, len, 1.0 / beta, 0.0, s))) return rc;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int j = 0;
    for (; j < m; ++j) {
      double* vj = k.V + (size_t)j * ld;
      double* vn = k.V + (size_t)(j + 1) * ld;
      if ((rc = A(vj, vn))) return rc;
      ++total;
      // CGS2: h = V^T w, w -= V h, twice (projections accumulated)
      if ((rc = dots(k, k.V, vn, len, j + 1, 0, k.h, s))) return rc;
      hipLaunchKernelGGL(k_combine, dim3(grid_for(len)), dim3(256), 0, s, vn, k.V, k.h, len, ld,
                         j + 1, 1.0, -1.0);
      if ((rc = hip_check(hipGetLastError(), "k_combine"))) return rc;
      if ((rc = dots(k, k.V, vn, len, j + 1, 0, proj2(k), s))) return rc;
      hipLaunchKernelGGL(k_combine, dim3(grid_for(len)), dim3(256), 0, s, vn, k.V, proj2(k), len,
                         ld, j + 1, 1.0, -1.0);
      if ((rc = hip_check(hipGetLastError(), "k_combine"))) return rc;
      // both projection sets in one copy (h holds them back to back)
      SDDC_HIP(hipMemcpyAsync(k.hpin, k.h, 2 * (size_t)(m + 2) * sizeof(double),
                              hipMemcpyDeviceToHost, s));
      double hn;
      if ((rc = norm2(k, vn, len, &hn, s))) return rc;   // synchronises the copy too
      for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = k.hpin[i] + k.hpin[(m + 2) + i];
      H[(size_t)(j + 1) * m + j] = hn;
      if (hn > 0.0 && (rc = axpby(vn, vn, nullptr, len, 1.0 / hn, 0.0, s))) return rc;
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
    std::copy(y.begin(), y.begin() + j, k.hpin);
    SDDC_HIP(hipMemcpyAsync(k.h, k.hpin, j * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_combine, dim3(grid_for(len)), dim3(256), 0, s, x, k.V, k.h, len, ld, j,
                       1.0, 1.0);
    if ((rc = hip_check(hipGetLastError(), "k_combine"))) return rc;
    SDDC_HIP(hipStreamSynchronize(s));   // the copy reads the pinned buffer
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

// out = mask(x) for a symmetric plan, else a copy.
static int masked_copy(sddc_krylov& k, const double* x, double* out, hipStream_t s) {
  if (k.plan->p.symmetric) return launch_mask(k.plan->p, 1, x, out, s);
  return axpby(out, x, nullptr, k.n, 1.0, 0.0, s);
}

// dfmu = DF(X, mu)_mu = [A4_BSub(dt Pr gr_k T), 0, 0] (PDFmu, Main.py:828-836):
// the psi part of one linear step of (0, T, 0) at Ra_T = 1, Ra_S = 0.
static int pdfmu(sddc_krylov& k, const double* X, double* dfmu, hipStream_t s) {
  const int N = k.n / 3;
  int rc;
  if ((rc = axpby(k.r, X, nullptr, k.n, 0.0, 0.0, s))) return rc;
  if ((rc = axpby(k.r + N, X + N, nullptr, N, 1.0, 0.0, s))) return rc;
  if ((rc = sddc_step(k.plan, 1, k.scal + 0, k.scal + 1, k.r, dfmu, 1, 1, s))) return rc;
  return axpby(dfmu + N, dfmu, nullptr, 2 * N, 0.0, 0.0, s);
}

}  // namespace sddc

using namespace sddc;

extern "C" {

int sddc_krylov_create(sddc_plan* plan, int restart, sddc_krylov** out) {
  if (!plan || !out) return fail(SDDC_ERR_INVALID, "null argument");
  if (restart < 1 || restart > 2000) return fail(SDDC_ERR_INVALID, "restart must be in [1, 2000]");
  *out = nullptr;
  std::lock_guard<std::recursive_mutex> lk(device_state_mutex());
  sddc_krylov* k = new sddc_krylov();
  k->plan = plan;
  k->n = 3 * plan->p.K * plan->p.nr;
  k->m = restart;
  const size_t n1 = (size_t)k->n + 1;
  auto A = [&](double** p, size_t count) {
    return hip_check(hipMalloc((void**)p, count * sizeof(double)), "krylov workspace");
  };
  int rc = A(&k->V, (size_t)(restart + 1) * basis_ld(k->n + 1));
  if (!rc) rc = A(&k->w, n1);
  if (!rc) rc = A(&k->r, n1);
  if (!rc) rc = A(&k->f, n1);
  if (!rc) rc = A(&k->t, n1);
  for (int i = 0; i < 5 && !rc; ++i) rc = A(&k->c[i], n1);
  if (!rc) rc = A(&k->scal, 4);
  if (!rc) rc = A(&k->partial, (size_t)(restart + 2) * kBlocks);
  if (!rc) rc = A(&k->h, 2 * ((size_t)restart + 2) + 1);
  if (!rc)
    rc = hip_check(hipHostMalloc((void**)&k->hpin, (2 * ((size_t)restart + 2) + 4) * sizeof(double),
                                 hipHostMallocDefault),
                   "krylov pinned host buffer");
  if (!rc) {
    k->hpin[0] = 1.0;
    k->hpin[1] = 0.0;
    rc = hip_check(hipMemcpy(k->scal, k->hpin, 2 * sizeof(double), hipMemcpyHostToDevice),
                   "krylov constants");
  }
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
  std::lock_guard<std::recursive_mutex> lk(device_state_mutex());
  double* bufs[] = {k->V, k->w, k->r, k->f, k->t, k->c[0], k->c[1], k->c[2], k->c[3], k->c[4],
                    k->scal, k->partial, k->h};
  for (double* b : bufs)
    if (b) (void)hipFree(b);
  if (k->hpin) (void)hipHostFree(k->hpin);
  delete k;
  return SDDC_OK;
}

int sddc_gmres(sddc_krylov* k, const double* Ra, const double* Ra_s, const double* X,
               const double* b, double* x, const sddc_krylov_opts* opts, sddc_krylov_info* info,
               void* stream) {
  if (!k || !X || !b || !x || !opts) return fail(SDDC_ERR_INVALID, "null argument");
  hipStream_t s = (hipStream_t)stream;
  const int fd = opts->finite_difference;
  int rc;
  if (fd && (rc = sddc_step(k->plan, 1, Ra, Ra_s, X, k->t, 1, 0, s))) return rc;
  const ApplyFn J = [&](const double* v, double* out) { return jv(*k, Ra, Ra_s, X, v, out, fd, s); };
  return gmres_core(*k, k->n, J, b, x, *opts, info, s);
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
    if ((rc = norm2(*k, k->f, n, &fnorm, s))) return rc;
    go.atol = opts->tol_gmres * fnorm;   // atol = tol_gmres * ||b|| (Main.py:534)
    if ((rc = axpby(k->w, k->f, nullptr, n, 0.0, 0.0, s))) return rc;   // dv = 0
    sddc_krylov_info gi{};
    if ((rc = sddc_gmres(k, Ra, Ra_s, X, k->f, k->w, &go, &gi, s))) return rc;
    krylov_total += gi.iterations;
    gmres_ok = gi.converged;   // the last solve's status decides, as in Main.py:541
    if ((rc = axpby(X, X, k->w, n, 1.0, -1.0, s))) return rc;
    double dn, xn;
    if ((rc = norm2(*k, k->w, n, &dn, s))) return rc;
    if ((rc = norm2(*k, X, n, &xn, s))) return rc;
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

int sddc_arclength_step(sddc_krylov* k, const double* Ra_s, double* Y, double* Y_dot,
                        double sign, double* ds, const sddc_krylov_opts* opts,
                        sddc_arclength_info* info, void* stream) {
  if (!k || !Ra_s || !Y || !Y_dot || !ds || !opts || !info)
    return fail(SDDC_ERR_INVALID, "null argument");
  hipStream_t s = (hipStream_t)stream;
  const int n = k->n, fd = opts->finite_difference;
  const double delta = 1.0 / n;   // delta = 1/(3N) (Main.py:777)
  const double tol = opts->tol_newton;
  double* X0 = k->c[0];     // masked start state X_0
  double* Xdot = k->c[1];   // X_dot
  double* dfmu = k->c[2];   // DF_mu at the linearisation point
  double* Xc = k->c[3];     // masked state of the current corrector iteration
  double* dY = k->c[4];     // corrector update / predictor xi
  double* mu_c = k->scal + 2;   // Ra of the linearisation point (device)
  double* dot = k->scal + 3;
  *info = sddc_arclength_info{};
  int rc, krylov_total = 0;
  sddc_krylov_opts go = *opts;

  // Bordered operator DG(dY) at (Xc, mu_c) (Main.py:912-913)
  double mudot = 0.0;
  const ApplyFn J = [&](const double* v, double* out) { return jv(*k, mu_c, Ra_s, Xc, v, out, fd, s); };
  const ApplyFn DG = [&](const double* v, double* out) {
    int r;
    if ((r = jv(*k, mu_c, Ra_s, Xc, v, out, fd, s))) return r;
    if ((r = dots(*k, Xdot, v, n, 1, 0, dot, s))) return r;
    hipLaunchKernelGGL(k_border, dim3(grid_for(n)), dim3(256), 0, s, out, v, dfmu, dot, n, delta,
                       mudot);
    return hip_check(hipGetLastError(), "k_border");
  };

  // Predict (Main.py:838-873)
  if ((rc = masked_copy(*k, Y, X0, s))) return rc;
  double mu0, ydn;
  if ((rc = get_scalar(*k, Y + n, &mu0, s))) return rc;
  if ((rc = norm2(*k, Y_dot, n + 1, &ydn, s))) return rc;
  if (ydn == 0.0 || k->plan->p.dt < 10.0) {
    // xi solves DF_X(X_0) xi = -DF_mu(X_0)
    if ((rc = set_scalar(mu_c, mu0, s))) return rc;
    if ((rc = axpby(Xc, X0, nullptr, n, 1.0, 0.0, s))) return rc;
    if ((rc = pdfmu(*k, Xc, k->f, s))) return rc;
    if ((rc = axpby(k->f, k->f, nullptr, n, -1.0, 0.0, s))) return rc;
    if (fd && (rc = sddc_step(k->plan, 1, mu_c, Ra_s, Xc, k->t, 1, 0, s))) return rc;
    double bn;
    if ((rc = norm2(*k, k->f, n, &bn, s))) return rc;
    go.atol = tol * bn;   // atol = tol_newton ||b|| (Main.py:857)
    if ((rc = axpby(dY, dY, nullptr, n, 0.0, 0.0, s))) return rc;
    sddc_krylov_info gi{};
    if ((rc = gmres_core(*k, n, J, k->f, dY, go, &gi, s))) return rc;
    krylov_total += gi.iterations;
    if (!gi.converged) {
      info->status = SDDC_ARC_PREDICTOR_FAILED;
      info->krylov_iterations = krylov_total;
      return SDDC_OK;
    }
    double nxi;
    if ((rc = norm2(*k, dY, n, &nxi, s))) return rc;
    mudot = sign / std::sqrt(1.0 + delta * (nxi - 1.0));   // as written in Main.py:862
    if ((rc = axpby(Xdot, dY, nullptr, n, mudot, 0.0, s))) return rc;
  } else {
    if ((rc = axpby(Xdot, Y_dot, nullptr, n, 1.0, 0.0, s))) return rc;
    if ((rc = get_scalar(*k, Y_dot + n, &mudot, s))) return rc;
  }
  double step = *ds;
  auto predict = [&]() {
    int r;
    if ((r = axpby(Y, X0, Xdot, n, 1.0, step, s))) return r;
    return set_scalar(Y + n, mu0 + mudot * step, s);
  };
  if ((rc = predict())) return rc;

  // Correct (Main.py:880-927)
  double err_X = 1.0, err_mu = 1.0;
  int it = 0;
  while (err_X > tol || err_mu > tol || it < 2) {
    if (it >= 5) {   // ds control (Main.py:883-896)
      it = 0;
      step *= 0.5;
      ++info->halvings;
      if (step < tol) {
        info->status = SDDC_ARC_STEP_TOO_SMALL;
        *ds = step;
        info->krylov_iterations = krylov_total;
        return SDDC_OK;
      }
      if ((rc = predict())) return rc;
    }
    double mu;
    if ((rc = masked_copy(*k, Y, Xc, s))) return rc;
    if ((rc = get_scalar(*k, Y + n, &mu, s))) return rc;
    if ((rc = set_scalar(mu_c, mu, s))) return rc;
    if ((rc = pdfmu(*k, Xc, dfmu, s))) return rc;
    // G = (Step(Xc) - Xc, delta X_dot.(Xc - X_0) + (1 - delta) mu_dot (mu - mu_0) - ds)
    if ((rc = sddc_step(k->plan, 1, mu_c, Ra_s, Xc, k->t, 1, 0, s))) return rc;
    if ((rc = axpby(k->f, k->t, Xc, n, 1.0, -1.0, s))) return rc;
    if ((rc = axpby(k->r, Xc, X0, n, 1.0, -1.0, s))) return rc;
    if ((rc = dots(*k, Xdot, k->r, n, 1, 0, dot, s))) return rc;
    double xdot_dx, gx2;
    if ((rc = get_scalar(*k, dot, &xdot_dx, s))) return rc;
    const double gmu = delta * xdot_dx + (1.0 - delta) * mudot * (mu - mu0) - step;
    if ((rc = set_scalar(k->f + n, gmu, s))) return rc;
    if ((rc = dots(*k, nullptr, k->f, n, 1, 0, dot, s))) return rc;
    if ((rc = get_scalar(*k, dot, &gx2, s))) return rc;
    const double bn = std::sqrt(delta * gx2 + (1.0 - delta) * gmu * gmu);
    go.atol = opts->tol_gmres * bn;   // Main.py:917-918
    if ((rc = axpby(dY, dY, nullptr, n + 1, 0.0, 0.0, s))) return rc;
    sddc_krylov_info gi{};
    if ((rc = gmres_core(*k, n + 1, DG, k->f, dY, go, &gi, s))) return rc;
    krylov_total += gi.iterations;
    if ((rc = axpby(Y, Y, dY, n + 1, 1.0, -1.0, s))) return rc;
    if (!gi.converged) {
      info->status = SDDC_ARC_CORRECTOR_FAILED;
      *ds = step;
      info->iterations = it;
      info->krylov_iterations = krylov_total;
      return SDDC_OK;
    }
    double ndx, nx, dmu;
    if ((rc = norm2(*k, dY, n, &ndx, s))) return rc;
    if ((rc = norm2(*k, Xc, n, &nx, s))) return rc;
    if ((rc = get_scalar(*k, dY + n, &dmu, s))) return rc;
    err_X = ndx / nx;
    err_mu = std::fabs(dmu) / std::fabs(mu);
    ++it;
  }
  if (it <= 4) step *= 2.0;   // Main.py:930-932
  *ds = step;
  info->iterations = it;
  info->error_X = err_X;
  info->error_mu = err_mu;

  // Tangent: DG Y_dot = e_last with the last corrector's operator, then
  // normalised in the weighted norm (Main.py:941-950)
  if ((rc = axpby(k->f, k->f, nullptr, n + 1, 0.0, 0.0, s))) return rc;
  if ((rc = set_scalar(k->f + n, 1.0, s))) return rc;
  go.atol = tol;   // tol_newton * ||e_last||
  if ((rc = axpby(Y_dot, Y_dot, nullptr, n + 1, 0.0, 0.0, s))) return rc;
  sddc_krylov_info gi{};
  if ((rc = gmres_core(*k, n + 1, DG, k->f, Y_dot, go, &gi, s))) return rc;
  krylov_total += gi.iterations;
  info->krylov_iterations = krylov_total;
  if (!gi.converged) {   // Main.py:943-944: a zero tangent
    info->tangent_ok = 0;
    return axpby(Y_dot, Y_dot, nullptr, n + 1, 0.0, 0.0, s);
  }
  info->tangent_ok = 1;
  double t2, tmu;
  if ((rc = dots(*k, nullptr, Y_dot, n, 1, 0, dot, s))) return rc;
  if ((rc = get_scalar(*k, dot, &t2, s))) return rc;
  if ((rc = get_scalar(*k, Y_dot + n, &tmu, s))) return rc;
  const double scale = std::sqrt(delta * t2 + (1.0 - delta) * tmu * tmu);
  if ((rc = axpby(Y_dot, Y_dot, nullptr, n + 1, 1.0 / scale, 0.0, s))) return rc;
  return hip_check(hipStreamSynchronize(s), "arclength step");
}

}  // extern "C"
