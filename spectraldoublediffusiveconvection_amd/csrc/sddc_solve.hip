// Implicit solves of the step, float64 always: block back substitution over
// theta modes with pre-inverted per-mode radial blocks.
//   A4_BSub_TSTEP_V2   (Matrix_Operators.py:1115-1194)  field 0 (psi)
//   NAB2_BSub_TSTEP_V2 (Matrix_Operators.py:1033-1086)  fields 1 (T), 2 (S)
// The recurrence over modes is sequential; the two parity chains of each
// field are independent.  One wavefront handles one (field, chain, instance);
// lane i owns radial row i: its row of the next block inverse is fetched into
// registers one mode ahead, and the right-hand side vector is broadcast
// through LDS.  The block inverses are stored transposed (column c of a block
// is contiguous), so each load instruction of a row fetch is one coalesced
// read across the wavefront.
#include "sddc_internal.h"

namespace sddc {

struct SolveArgs {
  const double* G;
  double* X;
  double* Pout;
  const double* LinvA4;
  const double* LinvT;
  const double* LinvS;
  const double* D2A4;
  const double* vecs;
  long long gstride, xstride;
  long long goff[3], xoff[3];
  int K, nr, NRP, symmetric, fields_mask;
  double dt, Pr, Tau;
};

// The solve workgroup is one wavefront: LDS writes of a wavefront are seen by
// its later LDS reads, so ordering needs only a compiler barrier, not
// __syncthreads() (whose s_waitcnt vmcnt(0) would also drain the operator
// rows prefetched for the next mode).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NRS>
__device__ __forceinline__ void load_row(double (&Lr)[NRS], const double* __restrict__ blockT,
                                         int NRP, int ii) {
#pragma unroll
  for (int c = 0; c < NRS; ++c) Lr[c] = blockT[(size_t)c * NRP + ii];
}

template <int NRS>
__device__ __forceinline__ double dot_row(const double (&Lr)[NRS], const double* shv) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
  for (int c = 0; c < NRS; c += 4) {
    s0 = fma(Lr[c], shv[c], s0);
    s1 = fma(Lr[c + 1], shv[c + 1], s1);
    s2 = fma(Lr[c + 2], shv[c + 2], s2);
    s3 = fma(Lr[c + 3], shv[c + 3], s3);
  }
  return (s0 + s1) + (s2 + s3);
}

// The steps below have no data-dependent branches: the prefetches of the next
// mode always issue (indices clamped at the last mode) and the first-row cases
// of the reference are selects.  With straight-line loop bodies the compiler
// can wait only for the loads of the previous step instead of draining all
// outstanding loads (s_waitcnt vmcnt(0)) every step.

// One NAB2 step (Matrix_Operators.py:1061-1083): b += 2 dt (j+2) f_{j+2};
// f_j = Linv_{K-1-j} (g_j - c_j b), c_0 = 1/2.  At the first row of a chain
// fprev = 0, so the unconditional update adds a signed zero (exact).
template <int NRS>
__device__ __forceinline__ void nab2_step(int& j, int K, int nr, int i, bool active, double dtf,
                                          const double* __restrict__ g, double* __restrict__ x,
                                          const double* __restrict__ Linv, size_t MS, int NRP,
                                          int ii, double (&Lc)[NRS], double (&Ln)[NRS],
                                          double& gcur, double& bacc, double& fprev, double* sh) {
  const int jj = K - 1 - j;
  const int jn = jj + 2 < K ? jj + 2 : K - 1;
  load_row<NRS>(Ln, Linv + (size_t)jn * MS, NRP, ii);
  const double gv = gcur;
  gcur = g[(size_t)(j >= 2 ? j - 2 : 0) * nr + ii];
  double bk = -2.0 * (j + 2.0);
  bk = -dtf * bk;
  bacc += bk * fprev;
  const double rhs = (j == 0) ? gv - 0.5 * bacc : gv - bacc;
  sh[i] = active ? rhs : 0.0;
  wave_sync();
  const double y = dot_row<NRS>(Lc, sh);
  wave_sync();
  if (active) x[(size_t)j * nr + i] = y;
  fprev = y;
  j -= 2;
}

// One A4 step (Matrix_Operators.py:1141-1164), dt = Pr*dt:
//   f_e += f_{row+2};  b = dt bjt (L1 f_e + IR4 bf_e) - bjt IR2 f_e
//   f_row = Linv (g_row + b);  bf_e += bj f_row + bjt f_e
// At the first row of a chain f_e = bf_e = fprev = 0, so b is a signed zero
// and the updates reduce exactly to the reference's first-row branch.
template <int NRS>
__device__ __forceinline__ void a4_step(int& row, int K, int nr, int i, bool active, double pdt,
                                        const double* __restrict__ g, double* __restrict__ x,
                                        double* __restrict__ Pp, const double* __restrict__ Linv,
                                        size_t MS, int NRP, int ii, double (&Lc)[NRS],
                                        double (&Ln)[NRS], const double* sD2T, double d2ii,
                                        double ir4, double ir2, double& gcur, double& fe,
                                        double& bf, double& fprev, double* sh) {
  const int jj = K - 1 - row;
  const int jn = jj + 2 < K ? jj + 2 : K - 1;
  load_row<NRS>(Ln, Linv + (size_t)jn * MS, NRP, ii);
  const double gv = gcur;
  gcur = g[(size_t)(row >= 2 ? row - 2 : 0) * nr + ii];
  const double j = (double)(K - jj);
  const double bj = -j * (j + 1.0);
  const double bjt = -2.0 * j;
  fe += fprev;
  sh[i] = active ? fe : 0.0;
  wave_sync();
  const double diag = d2ii + bj * ir4;   // L1 = D2 + bj IR4 (IR4 diagonal)
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int c = 0; c < NRS; c += 2) {
    const double c0 = (c == i) ? diag : sD2T[c * NRS + ii];
    const double c1 = (c + 1 == i) ? diag : sD2T[(c + 1) * NRS + ii];
    s0 = fma(c0, sh[c], s0);
    s1 = fma(c1, sh[c + 1], s1);
  }
  const double L1f = s0 + s1;
  wave_sync();
  const double bb = (pdt * bjt) * (L1f + ir4 * bf) - bjt * (ir2 * fe);
  sh[i] = active ? gv + bb : 0.0;
  wave_sync();
  const double y = dot_row<NRS>(Lc, sh);
  wave_sync();
  bf += bj * y + bjt * fe;
  if (active) {
    if (Pp) Pp[(size_t)row * nr + i] = fe;
    x[(size_t)row * nr + i] = y;
  }
  fprev = y;
  row -= 2;
}

template <int NRS>
__global__ __launch_bounds__(64) void k_solve(SolveArgs a) {
  __shared__ double sh[64];
  __shared__ double sD2T[NRS * NRS];   // D2 of the A4 solve, transposed
  const int field = blockIdx.x >> 1, chain = blockIdx.x & 1;
  if (!((a.fields_mask >> field) & 1)) return;
  const int b = blockIdx.y, i = threadIdx.x;
  const int K = a.K, nr = a.nr, NRP = a.NRP;
  const bool active = i < nr;
  const int ii = active ? i : 0;
  const double* g = a.G + (size_t)b * a.gstride + a.goff[field];
  double* x = a.X + (size_t)b * a.xstride + a.xoff[field];
  double* Pp = (field == 0 && a.Pout) ? a.Pout + (size_t)b * K * nr : nullptr;
  const int start = (chain == 0) ? K - 1 : K - 2;
  if (start < 0) return;
  // Eq_SYM keeps odd psi rows (A4 chain from K-1) and even T/S rows (NAB2
  // chain from K-2); the other chain is not solved (Matrix_Operators.py:1054,1166).
  const bool skip = a.symmetric && (field == 0 ? chain == 1 : chain == 0);
  if (skip) {
    for (int r = start; r >= 0; r -= 2)
      if (active) {
        x[(size_t)r * nr + i] = 0.0;
        if (Pp) Pp[(size_t)r * nr + i] = 0.0;
      }
    return;
  }
  const size_t MS = (size_t)NRP * NRP;
  const int nsteps = start / 2 + 1;
  double La[NRS], Lb[NRS];
  double fprev = 0.0;
  double gcur = g[(size_t)start * nr + ii];
  if (field > 0) {
    const double* Linv = (field == 1) ? a.LinvT : a.LinvS;
    const double dtf = (field == 1) ? a.dt : a.Tau * a.dt;
    double bacc = 0.0;
    load_row<NRS>(La, Linv + (size_t)(K - 1 - start) * MS, NRP, ii);
    int j = start;
    for (int s = 0; s + 1 < nsteps; s += 2) {
      nab2_step<NRS>(j, K, nr, i, active, dtf, g, x, Linv, MS, NRP, ii, La, Lb, gcur, bacc, fprev, sh);
      nab2_step<NRS>(j, K, nr, i, active, dtf, g, x, Linv, MS, NRP, ii, Lb, La, gcur, bacc, fprev, sh);
    }
    if (nsteps & 1)
      nab2_step<NRS>(j, K, nr, i, active, dtf, g, x, Linv, MS, NRP, ii, La, Lb, gcur, bacc, fprev, sh);
    return;
  }
  for (int e = threadIdx.x; e < NRS * NRS; e += blockDim.x) {
    const int r = e / NRS, c = e - r * NRS;
    sD2T[c * NRS + r] = a.D2A4[(size_t)r * NRP + c];
  }
  __syncthreads();
  const double ir4 = a.vecs[V_IR4A4 * NRP + ii];
  const double ir2 = a.vecs[V_IR2A4 * NRP + ii];
  const double d2ii = sD2T[ii * NRS + ii];
  const double pdt = a.Pr * a.dt;
  double fe = 0.0, bf = 0.0;
  load_row<NRS>(La, a.LinvA4 + (size_t)(K - 1 - start) * MS, NRP, ii);
  int row = start;
  for (int s = 0; s + 1 < nsteps; s += 2) {
    a4_step<NRS>(row, K, nr, i, active, pdt, g, x, Pp, a.LinvA4, MS, NRP, ii, La, Lb, sD2T, d2ii,
                 ir4, ir2, gcur, fe, bf, fprev, sh);
    a4_step<NRS>(row, K, nr, i, active, pdt, g, x, Pp, a.LinvA4, MS, NRP, ii, Lb, La, sD2T, d2ii,
                 ir4, ir2, gcur, fe, bf, fprev, sh);
  }
  if (nsteps & 1)
    a4_step<NRS>(row, K, nr, i, active, pdt, g, x, Pp, a.LinvA4, MS, NRP, ii, La, Lb, sD2T, d2ii,
                 ir4, ir2, gcur, fe, bf, fprev, sh);
}

int launch_solve_ex(const Plan& p, int B, const double* G, long long gstride,
                    const long long goff[3], double* X, long long xstride,
                    const long long xoff[3], double* Pout, int fields_mask, hipStream_t s) {
  if (B <= 0) return SDDC_OK;
  SolveArgs a;
  a.G = G;
  a.X = X;
  a.Pout = Pout;
  a.LinvA4 = p.LinvA4;
  a.LinvT = p.LinvT;
  a.LinvS = p.LinvS;
  a.D2A4 = p.D2A4;
  a.vecs = p.vecs;
  a.gstride = gstride;
  a.xstride = xstride;
  for (int f = 0; f < 3; ++f) {
    a.goff[f] = goff[f];
    a.xoff[f] = xoff[f];
  }
  a.K = p.K;
  a.nr = p.nr;
  a.NRP = p.NRP;
  a.symmetric = p.symmetric;
  a.fields_mask = fields_mask;
  a.dt = p.dt;
  a.Pr = p.Pr;
  a.Tau = p.Tau;
  const dim3 grid(6, B);
  const int nrs = ((p.nr + 7) / 8) * 8;
  switch (nrs) {
    case 8: hipLaunchKernelGGL(k_solve<8>, grid, dim3(64), 0, s, a); break;
    case 16: hipLaunchKernelGGL(k_solve<16>, grid, dim3(64), 0, s, a); break;
    case 24: hipLaunchKernelGGL(k_solve<24>, grid, dim3(64), 0, s, a); break;
    case 32: hipLaunchKernelGGL(k_solve<32>, grid, dim3(64), 0, s, a); break;
    case 40: hipLaunchKernelGGL(k_solve<40>, grid, dim3(64), 0, s, a); break;
    case 48: hipLaunchKernelGGL(k_solve<48>, grid, dim3(64), 0, s, a); break;
    case 56: hipLaunchKernelGGL(k_solve<56>, grid, dim3(64), 0, s, a); break;
    case 64: hipLaunchKernelGGL(k_solve<64>, grid, dim3(64), 0, s, a); break;
    default: return fail(SDDC_ERR_UNSUPPORTED, "unsupported radial size");
  }
  return hip_check(hipGetLastError(), "k_solve");
}

int launch_solve(const Plan& p, int B, const double* G, double* Xout, double* Pout,
                 int fields_mask, int /*mask_out*/, hipStream_t s) {
  const long long N = (long long)p.K * p.nr;
  const long long off[3] = {0, N, 2 * N};
  return launch_solve_ex(p, B, G, 3 * N, off, Xout, 3 * N, off, Pout, fields_mask, s);
}

}  // namespace sddc
