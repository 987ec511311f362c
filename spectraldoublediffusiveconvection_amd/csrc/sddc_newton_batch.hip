// Batched matrix-free Newton-Krylov (_Newton, Main.py:430-554) for B
// instances of one size class at once, float64 throughout.
//
// Every instance runs the same algorithm as the single-instance solver
// (sddc_newton): residual F(X) = Step(X) - X, GMRES(m) with classical
// Gram-Schmidt applied twice on J v = PDFX(v) = Step'(X) v - v (or a finite
// difference of the step), X <- X - dv, converged when ||dv|| / ||X|| <= tol
// within max_newton iterations and the last Krylov solve converged.  The B
// instances advance in lockstep: one J v is one launch of the batched step
// kernels for all of them, and every vector operation runs over all B with a
// per-instance gate.  The small least-squares problem of each instance
// (Hessenberg column, Givens rotations, residual estimate, back substitution)
// is updated on the device, one thread per instance, so an Arnoldi step needs
// no copy to the host; the host reads the instances' activity flags (B ints,
// pinned memory) only every `check_every` Arnoldi steps, at each restart and
// once per Newton iteration.  An instance that has converged is frozen (its
// least-squares data and solution are no longer touched) while the others go
// on, so its result does not depend on the other instances of the batch: the
// dot products reduce over a fixed per-instance split into kBlocks partial
// sums and a fixed-order sum, and the step kernels' launch shapes are the
// plan's.  A batch of one is how a single job runs, so a job gives the same
// bits alone and batched with others.
//
// Basis layout: V[j][b][n] (j = 0..m, b < capacity, n = 3N): the j-th basis
// vectors of all instances are one contiguous batch of states for the step
// kernels.
#include <cmath>
#include <vector>

#include "sddc_internal.h"

namespace sddc {
namespace {

constexpr int kThreads = 256;
constexpr int kBlocks = 240;   // partial sums per instance and dot product

// per-instance scalars (double) and flags (int)
enum Scal { S_NORM2 = 0, S_TARGET, S_INV, S_EPS, S_RES, S_XNORM2, S_VNORM2, S_ERR, S_BNORM, kNS };
enum Flag { F_ACTIVE = 0, F_GCONV, F_NEWTON, F_JEND, F_ITERS, F_CYCLE, F_RESTARTS, F_ZEROV, kNF };

struct Ctx {
  int n, m, cap, B;
  double *V, *F, *dv, *R, *T, *partial, *h, *h2, *H, *cs, *sn, *g, *y, *sc;
  int* fl;
};

// partial[b][i][blk] = (V_i[b] . x[b]) over this block's chunk, i < nv; V == null: x . x
__global__ __launch_bounds__(kThreads) void k_bdots(const double* __restrict__ V, size_t strideJ,
                                                    const double* __restrict__ x, int n, int nv,
                                                    int pstride, double* __restrict__ partial,
                                                    const int* __restrict__ fl, int gate) {
  const int b = blockIdx.y;
  if (gate >= 0 && !fl[(size_t)b * kNF + gate]) return;
  __shared__ double red[kThreads];
  const int chunk = (n + gridDim.x - 1) / gridDim.x;
  const int e0 = blockIdx.x * chunk;
  const int e1 = min(n, e0 + chunk);
  const double* xb = x + (size_t)b * n;
  for (int i = 0; i < nv; ++i) {
    const double* a = V ? V + i * strideJ + (size_t)b * n : xb;
    double s = 0.0;
    for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) s = fma(a[e], xb[e], s);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = kThreads / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[((size_t)b * pstride + i) * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

// out[b*ostride + i] = sum over blocks, fixed order
__global__ void k_breduce(const double* __restrict__ partial, int pstride, int nv,
                          double* __restrict__ out, int ostride, const int* __restrict__ fl,
                          int gate) {
  const int b = blockIdx.x;
  if (gate >= 0 && !fl[(size_t)b * kNF + gate]) return;
  for (int i = threadIdx.x; i < nv; i += blockDim.x) {
    const double* p = partial + ((size_t)b * pstride + i) * kBlocks;
    double s = 0.0;
    for (int q = 0; q < kBlocks; ++q) s += p[q];
    out[(size_t)b * ostride + i] = s;
  }
}

// y[b] = alpha y[b] + beta sum_{i < nv_b} c[b][i] V_i[b]  (nv_b = the instance's
// F_JEND when nv < 0, else nv)
__global__ void k_bcombine(double* __restrict__ y, const double* __restrict__ V, size_t strideJ,
                           const double* __restrict__ c, int cstride, int nv, int n, double alpha,
                           double beta, const int* __restrict__ fl, int gate) {
  const int b = blockIdx.y;
  if (gate >= 0 && !fl[(size_t)b * kNF + gate]) return;
  const int nvb = nv < 0 ? fl[(size_t)b * kNF + F_JEND] : nv;
  const double* cb = c + (size_t)b * cstride;
  double* yb = y + (size_t)b * n;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int i = 0; i < nvb; ++i) s = fma(cb[i], V[i * strideJ + (size_t)b * n + e], s);
    yb[e] = alpha * yb[e] + beta * s;
  }
}

// z[b] = a x[b] + b y[b] for gated instances (y may be null).  z may be x or
// y (in-place updates), so the three are not declared restrict.
__global__ void k_baxpby(double* z, const double* x, const double* y, int n, double a, double c,
                         const int* __restrict__ fl, int gate) {
  const int b = blockIdx.y;
  if (gate >= 0 && !fl[(size_t)b * kNF + gate]) return;
  const size_t o = (size_t)b * n;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    z[o + e] = a * x[o + e] + (y ? c * y[o + e] : 0.0);
}

// v[b] *= sc[b][S_INV]
__global__ void k_bscale(double* __restrict__ v, int n, const double* __restrict__ sc,
                         const int* __restrict__ fl, int gate) {
  const int b = blockIdx.y;
  if (gate >= 0 && !fl[(size_t)b * kNF + gate]) return;
  const double s = sc[(size_t)b * kNS + S_INV];
  const size_t o = (size_t)b * n;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    v[o + e] *= s;
}

// Start of a GMRES cycle for instances that still need Krylov work
// (Newton-active and not yet converged): beta = ||r||; converged when
// beta <= target, else g = beta e_1 and V_0 = r / beta.
__global__ void k_cycle_start(double* __restrict__ sc, int* __restrict__ fl,
                              double* __restrict__ g, int m, int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double* s = sc + (size_t)b * kNS;
  int* f = fl + (size_t)b * kNF;
  f[F_ACTIVE] = 0;
  f[F_CYCLE] = 0;
  if (!f[F_NEWTON] || f[F_GCONV]) return;
  const double beta = sqrt(s[S_NORM2]);
  s[S_RES] = beta;
  if (beta <= s[S_TARGET]) {
    f[F_GCONV] = 1;
    return;
  }
  double* gb = g + (size_t)b * (m + 1);
  for (int i = 0; i <= m; ++i) gb[i] = 0.0;
  gb[0] = beta;
  s[S_INV] = 1.0 / beta;
  f[F_ACTIVE] = 1;
  f[F_CYCLE] = 1;
  f[F_JEND] = 0;
}

// Arnoldi step j of every active instance: Hessenberg column j from the two
// projection sets and ||w||, the previous rotations, a new rotation, the
// residual estimate |g_{j+1}|; inactive from now on when it is <= target or
// w vanished.  S_INV gets 1/||w|| for the normalisation of the new vector.
__global__ void k_givens(const double* __restrict__ h, const double* __restrict__ h2,
                         double* __restrict__ H, double* __restrict__ cs, double* __restrict__ sn,
                         double* __restrict__ g, double* __restrict__ sc, int* __restrict__ fl,
                         int j, int m, int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  int* f = fl + (size_t)b * kNF;
  if (!f[F_ACTIVE]) return;
  double* s = sc + (size_t)b * kNS;
  const double* hb = h + (size_t)b * (m + 2);
  const double* h2b = h2 + (size_t)b * (m + 2);
  double* Hb = H + (size_t)b * (m + 1) * m;
  double* csb = cs + (size_t)b * m;
  double* snb = sn + (size_t)b * m;
  double* gb = g + (size_t)b * (m + 1);
  const double hn = sqrt(s[S_NORM2]);
  for (int i = 0; i <= j; ++i) Hb[(size_t)i * m + j] = hb[i] + h2b[i];
  Hb[(size_t)(j + 1) * m + j] = hn;
  for (int i = 0; i < j; ++i) {
    const double a = Hb[(size_t)i * m + j], c = Hb[(size_t)(i + 1) * m + j];
    Hb[(size_t)i * m + j] = csb[i] * a + snb[i] * c;
    Hb[(size_t)(i + 1) * m + j] = -snb[i] * a + csb[i] * c;
  }
  const double a = Hb[(size_t)j * m + j], c = Hb[(size_t)(j + 1) * m + j];
  const double den = hypot(a, c);
  csb[j] = den > 0.0 ? a / den : 1.0;
  snb[j] = den > 0.0 ? c / den : 0.0;
  Hb[(size_t)j * m + j] = den;
  Hb[(size_t)(j + 1) * m + j] = 0.0;
  gb[j + 1] = -snb[j] * gb[j];
  gb[j] = csb[j] * gb[j];
  const double res = fabs(gb[j + 1]);
  s[S_RES] = res;
  s[S_INV] = hn > 0.0 ? 1.0 / hn : 0.0;
  f[F_JEND] = j + 1;
  f[F_ITERS] += 1;
  if (res <= s[S_TARGET] || hn == 0.0) f[F_ACTIVE] = 0;
}

// End of a cycle: y = H(0:jend, 0:jend)^-1 g(0:jend) for instances that ran
// it; converged when the residual estimate reached the target.
__global__ void k_backsolve(const double* __restrict__ H, const double* __restrict__ g,
                            double* __restrict__ y, double* __restrict__ sc,
                            int* __restrict__ fl, int m, int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  int* f = fl + (size_t)b * kNF;
  if (!f[F_CYCLE]) return;
  const double* Hb = H + (size_t)b * (m + 1) * m;
  const double* gb = g + (size_t)b * (m + 1);
  double* yb = y + (size_t)b * m;
  const int J = f[F_JEND];
  for (int i = J - 1; i >= 0; --i) {
    double acc = gb[i];
    for (int q = i + 1; q < J; ++q) acc -= Hb[(size_t)i * m + q] * yb[q];
    yb[i] = acc / Hb[(size_t)i * m + i];
  }
  f[F_RESTARTS] += 1;
  f[F_ACTIVE] = 0;
  if (sc[(size_t)b * kNS + S_RES] <= sc[(size_t)b * kNS + S_TARGET]) f[F_GCONV] = 1;
}

// Finite-difference point: eps = sqrt(eps_mach) (1 + ||X||) / ||v||,
// R = X + eps v (Newton-active instances).
__global__ void k_fd_point(double* __restrict__ R, const double* __restrict__ X,
                           const double* __restrict__ v, int n, double* __restrict__ sc,
                           int* __restrict__ fl) {
  const int b = blockIdx.y;
  int* f = fl + (size_t)b * kNF;
  if (!f[F_NEWTON]) return;
  double* s = sc + (size_t)b * kNS;
  const double nv = sqrt(s[S_VNORM2]);
  const double eps = nv > 0.0 ? 1.4901161193847656e-08 * (1.0 + sqrt(s[S_XNORM2])) / nv : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s[S_EPS] = eps;
    f[F_ZEROV] = nv > 0.0 ? 0 : 1;
  }
  const size_t o = (size_t)b * n;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    R[o + e] = X[o + e] + eps * v[o + e];
}

// out = (out - T) / eps - v, or 0 for a zero v
__global__ void k_fd_finish(double* __restrict__ out, const double* __restrict__ T,
                            const double* __restrict__ v, int n, const double* __restrict__ sc,
                            const int* __restrict__ fl) {
  const int b = blockIdx.y;
  const int* f = fl + (size_t)b * kNF;
  if (!f[F_NEWTON]) return;
  const double eps = sc[(size_t)b * kNS + S_EPS];
  const bool zero = f[F_ZEROV] != 0;
  const size_t o = (size_t)b * n;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    out[o + e] = zero ? 0.0 : (out[o + e] - T[o + e]) / eps - v[o + e];
}

// target = tol * ||F|| (atol of the Krylov solve, Main.py:534); ||b|| kept
__global__ void k_targets(double* __restrict__ sc, int* __restrict__ fl, double tol, int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double* s = sc + (size_t)b * kNS;
  int* f = fl + (size_t)b * kNF;
  s[S_BNORM] = sqrt(s[S_NORM2]);
  s[S_TARGET] = tol * s[S_BNORM];
  f[F_GCONV] = 0;
  f[F_ITERS] = 0;
  f[F_RESTARTS] = 0;
}

// err = ||dv|| / ||X|| (S_VNORM2 holds ||dv||^2, S_XNORM2 ||X||^2)
__global__ void k_newton_err(double* __restrict__ sc, const int* __restrict__ fl, int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || !fl[(size_t)b * kNF + F_NEWTON]) return;
  double* s = sc + (size_t)b * kNS;
  const double xn = sqrt(s[S_XNORM2]);
  s[S_ERR] = xn > 0.0 ? sqrt(s[S_VNORM2]) / xn : INFINITY;
}

// sc[b][slot] = sum over the fixed per-instance split of x[b] . x[b]
int norms2(Ctx& c, const double* x, int slot, int gate, hipStream_t s) {
  hipLaunchKernelGGL(k_bdots, dim3(kBlocks, c.B), dim3(kThreads), 0, s, nullptr, (size_t)0, x,
                     c.n, 1, c.m + 2, c.partial, c.fl, gate);
  if (int rc = hip_check(hipGetLastError(), "k_bdots")) return rc;
  hipLaunchKernelGGL(k_breduce, dim3(c.B), dim3(64), 0, s, c.partial, c.m + 2, 1, c.sc + slot,
                     (int)kNS, c.fl, gate);
  return hip_check(hipGetLastError(), "k_breduce");
}

int vgrid(int n) { return std::min(64, (n + kThreads - 1) / kThreads); }

int axpby(Ctx& c, double* z, const double* x, const double* y, double a, double b, int gate,
          hipStream_t s) {
  hipLaunchKernelGGL(k_baxpby, dim3(vgrid(c.n), c.B), dim3(kThreads), 0, s, z, x, y, c.n, a, b,
                     c.fl, gate);
  return hip_check(hipGetLastError(), "k_baxpby");
}

int pgrid(int B) { return (B + 63) / 64; }

}  // namespace
}  // namespace sddc

struct sddc_newton_batch {
  sddc_plan* plan = nullptr;
  int n = 0, m = 0, cap = 0;
  double *V = nullptr, *F = nullptr, *dv = nullptr, *R = nullptr, *T = nullptr;
  double *partial = nullptr, *h = nullptr, *h2 = nullptr, *H = nullptr, *cs = nullptr,
         *sn = nullptr, *g = nullptr, *y = nullptr, *sc = nullptr;
  int* fl = nullptr;
  int* hfl = nullptr;       // pinned host copy of the flags
  double* hsc = nullptr;    // pinned host copy of the scalars
};

using namespace sddc;

namespace {

Ctx ctx_of(sddc_newton_batch* k, int B) {
  Ctx c{};
  c.n = k->n;
  c.m = k->m;
  c.cap = k->cap;
  c.B = B;
  c.V = k->V;
  c.F = k->F;
  c.dv = k->dv;
  c.R = k->R;
  c.T = k->T;
  c.partial = k->partial;
  c.h = k->h;
  c.h2 = k->h2;
  c.H = k->H;
  c.cs = k->cs;
  c.sn = k->sn;
  c.g = k->g;
  c.y = k->y;
  c.sc = k->sc;
  c.fl = k->fl;
  return c;
}

// J v for every instance into out (B vectors of n)
int jv(sddc_newton_batch* k, Ctx& c, const double* Ra, const double* Ras, const double* X,
       const double* v, double* out, int fd, hipStream_t s) {
  int rc;
  if (!fd) {
    if ((rc = sddc_step_linearized(k->plan, c.B, Ra, Ras, v, X, out, s))) return rc;
    return axpby(c, out, out, v, 1.0, -1.0, -1, s);
  }
  if ((rc = norms2(c, v, S_VNORM2, F_NEWTON, s))) return rc;
  hipLaunchKernelGGL(k_fd_point, dim3(vgrid(c.n), c.B), dim3(kThreads), 0, s, c.R, X, v, c.n,
                     c.sc, c.fl);
  if ((rc = hip_check(hipGetLastError(), "k_fd_point"))) return rc;
  if ((rc = sddc_step(k->plan, c.B, Ra, Ras, c.R, out, 1, 0, s))) return rc;
  hipLaunchKernelGGL(k_fd_finish, dim3(vgrid(c.n), c.B), dim3(kThreads), 0, s, out, c.T, v, c.n,
                     c.sc, c.fl);
  return hip_check(hipGetLastError(), "k_fd_finish");
}

// flags of all instances to the host (pinned) and wait for them
int read_flags(sddc_newton_batch* k, int B, hipStream_t s) {
  SDDC_HIP(hipMemcpyAsync(k->hfl, k->fl, (size_t)B * kNF * sizeof(int), hipMemcpyDeviceToHost, s));
  SDDC_HIP(hipStreamSynchronize(s));
  return SDDC_OK;
}

bool any_flag(const sddc_newton_batch* k, int B, int which) {
  for (int b = 0; b < B; ++b)
    if (k->hfl[(size_t)b * kNF + which]) return true;
  return false;
}

// GMRES(m) on J dv = F for every Newton-active instance (dv = 0 on entry).
int gmres(sddc_newton_batch* k, Ctx& c, const double* Ra, const double* Ras, const double* X,
          int fd, int max_restarts, int check_every, hipStream_t s) {
  const int m = c.m;
  const size_t sJ = (size_t)c.cap * c.n;   // stride between basis vectors j and j+1
  int rc;
  for (int r = 0; r < std::max(1, max_restarts); ++r) {
    // r = F - J dv (with dv = 0 on the first cycle: r = F)
    if (r == 0) {
      if ((rc = axpby(c, c.V, c.F, nullptr, 1.0, 0.0, F_NEWTON, s))) return rc;
    } else {
      if ((rc = jv(k, c, Ra, Ras, X, c.dv, c.V, fd, s))) return rc;
      if ((rc = axpby(c, c.V, c.F, c.V, 1.0, -1.0, F_NEWTON, s))) return rc;
    }
    if ((rc = norms2(c, c.V, S_NORM2, F_NEWTON, s))) return rc;
    hipLaunchKernelGGL(k_cycle_start, dim3(pgrid(c.B)), dim3(64), 0, s, c.sc, c.fl, c.g, m, c.B);
    if ((rc = hip_check(hipGetLastError(), "k_cycle_start"))) return rc;
    if ((rc = read_flags(k, c.B, s))) return rc;
    if (!any_flag(k, c.B, F_ACTIVE)) break;
    // V_0 = r / beta for the instances that run this cycle
    hipLaunchKernelGGL(k_bscale, dim3(vgrid(c.n), c.B), dim3(kThreads), 0, s, c.V, c.n, c.sc,
                       c.fl, (int)F_CYCLE);
    if ((rc = hip_check(hipGetLastError(), "k_bscale"))) return rc;
    for (int j = 0; j < m; ++j) {
      double* vj = c.V + j * sJ;
      double* vn = c.V + (j + 1) * sJ;
      if ((rc = jv(k, c, Ra, Ras, X, vj, vn, fd, s))) return rc;
      // CGS2 on the active instances: h = V^T w, w -= V h, twice
      for (int pass = 0; pass < 2; ++pass) {
        double* hp = pass ? c.h2 : c.h;
        hipLaunchKernelGGL(k_bdots, dim3(kBlocks, c.B), dim3(kThreads), 0, s, c.V, sJ, vn, c.n,
                           j + 1, m + 2, c.partial, c.fl, (int)F_ACTIVE);
        if ((rc = hip_check(hipGetLastError(), "k_bdots"))) return rc;
        hipLaunchKernelGGL(k_breduce, dim3(c.B), dim3(64), 0, s, c.partial, m + 2, j + 1, hp,
                           m + 2, c.fl, (int)F_ACTIVE);
        if ((rc = hip_check(hipGetLastError(), "k_breduce"))) return rc;
        hipLaunchKernelGGL(k_bcombine, dim3(vgrid(c.n), c.B), dim3(kThreads), 0, s, vn, c.V, sJ,
                           hp, m + 2, j + 1, c.n, 1.0, -1.0, c.fl, (int)F_ACTIVE);
        if ((rc = hip_check(hipGetLastError(), "k_bcombine"))) return rc;
      }
      if ((rc = norms2(c, vn, S_NORM2, F_ACTIVE, s))) return rc;
      hipLaunchKernelGGL(k_givens, dim3(pgrid(c.B)), dim3(64), 0, s, c.h, c.h2, c.H, c.cs, c.sn,
                         c.g, c.sc, c.fl, j, m, c.B);
      if ((rc = hip_check(hipGetLastError(), "k_givens"))) return rc;
      // normalise the new vector of the instances still iterating (an instance
      // that converged at this step never reads it)
      hipLaunchKernelGGL(k_bscale, dim3(vgrid(c.n), c.B), dim3(kThreads), 0, s, vn, c.n, c.sc,
                         c.fl, (int)F_ACTIVE);
      if ((rc = hip_check(hipGetLastError(), "k_bscale"))) return rc;
      if ((j + 1) % check_every == 0 || j + 1 == m) {
        if ((rc = read_flags(k, c.B, s))) return rc;
        if (!any_flag(k, c.B, F_ACTIVE)) break;
      }
    }
    // y per instance, dv += V y over its own jend columns
    hipLaunchKernelGGL(k_backsolve, dim3(pgrid(c.B)), dim3(64), 0, s, c.H, c.g, c.y, c.sc, c.fl, m,
                       c.B);
    if ((rc = hip_check(hipGetLastError(), "k_backsolve"))) return rc;
    hipLaunchKernelGGL(k_bcombine, dim3(vgrid(c.n), c.B), dim3(kThreads), 0, s, c.dv, c.V, sJ,
                       c.y, m, -1, c.n, 1.0, 1.0, c.fl, (int)F_CYCLE);
    if ((rc = hip_check(hipGetLastError(), "k_bcombine"))) return rc;
    if ((rc = read_flags(k, c.B, s))) return rc;
    bool more = false;
    for (int b = 0; b < c.B; ++b)
      more = more || (k->hfl[(size_t)b * kNF + F_NEWTON] && !k->hfl[(size_t)b * kNF + F_GCONV]);
    if (!more) break;
  }
  return SDDC_OK;
}

}  // namespace

extern "C" {

int sddc_newton_batch_create(sddc_plan* plan, int restart, int max_batch,
                             sddc_newton_batch** out) {
  if (!plan || !out) return fail(SDDC_ERR_INVALID, "null argument");
  if (restart < 1 || restart > 1000) return fail(SDDC_ERR_INVALID, "restart must be in [1, 1000]");
  if (max_batch < 1) return fail(SDDC_ERR_INVALID, "max_batch must be positive");
  *out = nullptr;
  std::lock_guard<std::recursive_mutex> lk(device_state_mutex());
  sddc_newton_batch* k = new sddc_newton_batch();
  k->plan = plan;
  k->n = 3 * plan->p.K * plan->p.nr;
  k->m = restart;
  k->cap = max_batch;
  const size_t n = k->n, m = restart, B = max_batch;
  auto A = [&](double** p, size_t count) {
    return hip_check(hipMalloc((void**)p, count * sizeof(double)), "newton batch workspace");
  };
  int rc = A(&k->V, (m + 1) * B * n);
  if (!rc) rc = A(&k->F, B * n);
  if (!rc) rc = A(&k->dv, B * n);
  if (!rc) rc = A(&k->R, B * n);
  if (!rc) rc = A(&k->T, B * n);
  if (!rc) rc = A(&k->partial, B * (m + 2) * kBlocks);
  if (!rc) rc = A(&k->h, B * (m + 2));
  if (!rc) rc = A(&k->h2, B * (m + 2));
  if (!rc) rc = A(&k->H, B * (m + 1) * m);
  if (!rc) rc = A(&k->cs, B * m);
  if (!rc) rc = A(&k->sn, B * m);
  if (!rc) rc = A(&k->g, B * (m + 1));
  if (!rc) rc = A(&k->y, B * m);
  if (!rc) rc = A(&k->sc, B * kNS);
  if (!rc)
    rc = hip_check(hipMalloc((void**)&k->fl, B * kNF * sizeof(int)), "newton batch flags");
  if (!rc)
    rc = hip_check(hipHostMalloc((void**)&k->hfl, B * kNF * sizeof(int), hipHostMallocDefault),
                   "newton batch pinned flags");
  if (!rc)
    rc = hip_check(hipHostMalloc((void**)&k->hsc, B * kNS * sizeof(double), hipHostMallocDefault),
                   "newton batch pinned scalars");
  if (!rc) rc = sddc_plan_reserve(plan, max_batch);
  if (rc) {
    sddc_newton_batch_destroy(k);
    return rc;
  }
  *out = k;
  return SDDC_OK;
}

int sddc_newton_batch_destroy(sddc_newton_batch* k) {
  if (!k) return SDDC_OK;
  std::lock_guard<std::recursive_mutex> lk(device_state_mutex());
  double* bufs[] = {k->V, k->F, k->dv, k->R, k->T, k->partial, k->h, k->h2, k->H, k->cs,
                    k->sn, k->g, k->y, k->sc};
  for (double* b : bufs)
    if (b) (void)hipFree(b);
  if (k->fl) (void)hipFree(k->fl);
  if (k->hfl) (void)hipHostFree(k->hfl);
  if (k->hsc) (void)hipHostFree(k->hsc);
  delete k;
  return SDDC_OK;
}

int sddc_newton_batched(sddc_newton_batch* k, int B, const double* Ra, const double* Ra_s,
                      double* X, const sddc_krylov_opts* opts, sddc_newton_info* info,
                      void* stream) {
  if (!k || !Ra || !Ra_s || !X || !opts || !info) return fail(SDDC_ERR_INVALID, "null argument");
  if (B < 1 || B > k->cap) return fail(SDDC_ERR_INVALID, "batch outside [1, max_batch]");
  hipStream_t s = (hipStream_t)stream;
  Ctx c = ctx_of(k, B);
  const int max_newton = opts->max_newton > 0 ? opts->max_newton : 5;
  const int check_every = opts->check_every > 0 ? opts->check_every : 4;
  const int fd = opts->finite_difference;
  const Plan& p = k->plan->p;
  int rc;
  // host bookkeeping of every instance: Newton iterations, Krylov steps, status
  std::vector<int> it(B, 0), kry(B, 0), ok(B, 1);
  std::vector<double> err(B, 1.0);
  for (int b = 0; b < B; ++b) {
    for (int q = 0; q < kNF; ++q) k->hfl[(size_t)b * kNF + q] = 0;
    k->hfl[(size_t)b * kNF + F_NEWTON] = 1;
  }
  SDDC_HIP(hipMemcpyAsync(k->fl, k->hfl, (size_t)B * kNF * sizeof(int), hipMemcpyHostToDevice, s));
  SDDC_HIP(hipStreamSynchronize(s));   // the pinned flags are reused below
  for (int round = 0; round < max_newton; ++round) {
    bool any = false;
    for (int b = 0; b < B; ++b) any = any || k->hfl[(size_t)b * kNF + F_NEWTON];
    if (!any) break;
    // F(X) = Step(X) - X (the step applies the plan's symmetry mask)
    if (p.symmetric && (rc = launch_mask(p, B, X, X, s))) return rc;
    if ((rc = sddc_step(k->plan, B, Ra, Ra_s, X, c.F, 1, 0, s))) return rc;
    if ((rc = axpby(c, c.F, c.F, X, 1.0, -1.0, F_NEWTON, s))) return rc;
    if ((rc = norms2(c, c.F, S_NORM2, F_NEWTON, s))) return rc;
    hipLaunchKernelGGL(k_targets, dim3(pgrid(B)), dim3(64), 0, s, c.sc, c.fl, opts->tol_gmres, B);
    if ((rc = hip_check(hipGetLastError(), "k_targets"))) return rc;
    // dv = 0 (0 * F: dv's first contents are uninitialised memory)
    if ((rc = axpby(c, c.dv, c.F, nullptr, 0.0, 0.0, F_NEWTON, s))) return rc;
    if (fd) {   // Step(X) and ||X||^2 at the linearisation point
      if ((rc = sddc_step(k->plan, B, Ra, Ra_s, X, c.T, 1, 0, s))) return rc;
      if ((rc = norms2(c, X, S_XNORM2, F_NEWTON, s))) return rc;
    }
    if ((rc = gmres(k, c, Ra, Ra_s, X, fd, opts->max_restarts, check_every, s))) return rc;
    // X -= dv; err = ||dv|| / ||X|| for the Newton-active instances
    if ((rc = axpby(c, X, X, c.dv, 1.0, -1.0, F_NEWTON, s))) return rc;
    if ((rc = norms2(c, c.dv, S_VNORM2, F_NEWTON, s))) return rc;
    if ((rc = norms2(c, X, S_XNORM2, F_NEWTON, s))) return rc;
    hipLaunchKernelGGL(k_newton_err, dim3(pgrid(B)), dim3(64), 0, s, c.sc, c.fl, B);
    if ((rc = hip_check(hipGetLastError(), "k_newton_err"))) return rc;
    SDDC_HIP(hipMemcpyAsync(k->hsc, c.sc, (size_t)B * kNS * sizeof(double), hipMemcpyDeviceToHost, s));
    if ((rc = read_flags(k, B, s))) return rc;
    for (int b = 0; b < B; ++b) {
      int* f = k->hfl + (size_t)b * kNF;
      if (!f[F_NEWTON]) continue;
      it[b] += 1;
      kry[b] += f[F_ITERS];
      ok[b] = f[F_GCONV];   // the last solve's status decides (Main.py:541)
      err[b] = k->hsc[(size_t)b * kNS + S_ERR];
      if (!(err[b] > opts->tol_newton && it[b] < max_newton)) f[F_NEWTON] = 0;
    }
    SDDC_HIP(hipMemcpyAsync(k->fl, k->hfl, (size_t)B * kNF * sizeof(int), hipMemcpyHostToDevice, s));
    SDDC_HIP(hipStreamSynchronize(s));
  }
  for (int b = 0; b < B; ++b) {
    info[b].iterations = it[b];
    info[b].converged = (it[b] < max_newton && ok[b]) ? 1 : 0;
    info[b].error = err[b];
    info[b].krylov_iterations = kry[b];
  }
  return SDDC_OK;
}

}  // extern "C"
