// FFT-based kernels: generic transforms (Transforms.py), the fused nonlinear
// advection kernel (NLIN_FX / NLIN_DFX, Matrix_Operators.py:743-898) and the
// radial rows of the kinetic energy (Main.py:71-134).
#include <atomic>
#include <type_traits>

#include "sddc_fft.h"

namespace sddc {

// ---------------------------------------------------------------------------
// Generic transforms.  One workgroup transforms two rows with one complex FFT.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_transform(FftShape sh, const double2* tw_,
                                                   const double2* rot_, int kind, int rows,
                                                   int len_in, int n_out,
                                                   const double* __restrict__ in,
                                                   double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int red[16 * 2];
  using C = cplx<double>;
  C* bufA = reinterpret_cast<C*>(smem_raw);
  C* bufB = bufA + sh.L;
  const C* tw = reinterpret_cast<const C*>(tw_);
  const C* rot = reinterpret_cast<const C*>(rot_);
  const int L = sh.L;
  const int r0 = 2 * blockIdx.x;
  const int r1 = r0 + 1;
  const bool two = r1 < rows;
  const double* u = in + (size_t)r0 * len_in;
  const double* v = two ? in + (size_t)r1 * len_in : nullptr;
  if (kind == SDDC_IDCT || kind == SDDC_IDST) {
    const bool is_sin = (kind == SDDC_IDST);
    // IDCT with n < len_in keeps the first n coefficients (DCT-III truncation)
    const int K = len_in < L ? len_in : L;
    int mx[2] = {series_maxexp<double>(u, K, is_sin, 0), series_maxexp<double>(v, K, is_sin, 0)};
    block_max_exp<2>(mx, red);
    const double su = pow2_scale_e<double>(mx[0]), sv = pow2_scale_e<double>(mx[1]);
    syn_pack<double>(bufA, u, is_sin, 0, v, is_sin, 0, K, L, rot, su, sv);
    __syncthreads();
    C* res = fft_run<double>(bufA, bufB, 1, sh, tw, true);
    const double iu = 1.0 / su, iv = 1.0 / sv;
    for (int q = threadIdx.x; q < L; q += blockDim.x) {
      const int n = grid_index(q, L);
      const double sg = is_sin ? (double)grid_sign(q, L) : 1.0;
      out[(size_t)r0 * n_out + n] = res[q].x * iu * sg;
      if (two) out[(size_t)r1 * n_out + n] = res[q].y * iv * sg;
    }
  } else {
    const bool is_sin = (kind == SDDC_DST);
    int mx[2] = {-1000, -1000};
    for (int q = threadIdx.x; q < L; q += blockDim.x) {
      mx[0] = max(mx[0], bexp(u[q]));
      if (v) mx[1] = max(mx[1], bexp(v[q]));
    }
    block_max_exp<2>(mx, red);
    const double su = pow2_scale_e<double>(mx[0]), sv = pow2_scale_e<double>(mx[1]);
    for (int q = threadIdx.x; q < L; q += blockDim.x) {
      const int n = grid_index(q, L);
      const double sg = is_sin ? (double)grid_sign(q, L) : 1.0;
      bufA[q] = {su * u[n] * sg, v ? sv * v[n] * sg : 0.0};
    }
    __syncthreads();
    C* res = fft_run<double>(bufA, bufB, 1, sh, tw, false);
    const int nk = n_out < L ? n_out : L;
    const double invL = 1.0 / L;
    const double iu = 1.0 / su, iv = 1.0 / sv;
    for (int k = threadIdx.x; k < nk; k += blockDim.x) {
      double c1, c2;
      if (!is_sin) {
        ana_unpack<double>(res, L, k, rot, c1, c2);
        double a1 = (2.0 * (c1 * iu)) * invL, a2 = (2.0 * (c2 * iv)) * invL;
        if (k == 0) { a1 *= 0.5; a2 *= 0.5; }
        out[(size_t)r0 * nk + k] = a1;
        if (two) out[(size_t)r1 * nk + k] = a2;
      } else {
        double b1 = 0.0, b2 = 0.0;
        if (k > 0) {
          ana_unpack<double>(res, L, L - k, rot, c1, c2);
          b1 = (2.0 * (c1 * iu)) * invL;
          b2 = (2.0 * (c2 * iv)) * invL;
        }
        out[(size_t)r0 * nk + k] = b1;
        if (two) out[(size_t)r1 * nk + k] = b2;
      }
    }
  }
}

int launch_transform(const TableSet& t, int kind, int rows, int len_in, int n_out,
                     const double* in, double* out, hipStream_t s) {
  if (rows <= 0) return SDDC_OK;
  const size_t lds = 2 * (size_t)t.L * sizeof(double2);
  if (lds > 160 * 1024) return fail(SDDC_ERR_UNSUPPORTED, "transform length too large for LDS");
  static std::atomic<size_t> cur{0};
  if (int rc = ensure_dynamic_lds((const void*)k_transform, lds, cur)) return rc;
  hipLaunchKernelGGL(k_transform, dim3((rows + 1) / 2), dim3(256), lds, s, t.shape, t.tw64,
                     t.rot64, kind, rows, len_in, n_out, in, out);
  return hip_check(hipGetLastError(), "k_transform");
}

// ---------------------------------------------------------------------------
// Nonlinear advection: one workgroup per (radial row i, instance b).
// Slab rows (layout [b][7][nr][K]): 0 JT (cos), 1 Dpsi (sine, sinusoid
// index), 2 Omega (sine), 3 Dr T (cos), 4 Dr S (cos), 5 T (cos row), 6 S.
// Output rows (layout [b][4][nr][K]): 0 sine coefficients of JT*omega (the
// radial derivative Dr is applied afterwards in spectral space, where it
// commutes with the theta transform), 1 sine coefficients of
// kDpsi*omega + Dpsi*komega, 2 cosine coefficients of the T product,
// 3 cosine coefficients of the S product (Matrix_Operators.py:791-799).
// ---------------------------------------------------------------------------
template <typename R>
struct SlabRows {
  const R* s[7];
};

template <typename R>
__device__ __forceinline__ SlabRows<R> slab_rows(const R* base, int b, int i, int nr, int K) {
  SlabRows<R> r;
#pragma unroll
  for (int s = 0; s < 7; ++s) r.s[s] = base + ((size_t)(b * 7 + s) * nr + i) * K;
  return r;
}

// The nine grid fields as series of the slab rows, in pair order:
//  0 JT (cos)   1 omega (sine)   | pair 0
//  2 komega (cos, m*omega_m)     3 Dpsi (sine)     | pair 1
//  4 kDpsi (cos, m*Dpsi_m)       5 kT (sine, -m T_m) | pair 2
//  6 DT (cos)   7 kC (sine, -m S_m)                  | pair 3
//  8 DC (cos)                                        | pair 4
template <typename R>
__device__ __forceinline__ void field_src(const SlabRows<R>& sr, int f, const R*& row, bool& is_sin,
                                          int& mult) {
  switch (f) {
    case 0: row = sr.s[0]; is_sin = false; mult = 0; break;
    case 1: row = sr.s[2]; is_sin = true; mult = 0; break;
    case 2: row = sr.s[2]; is_sin = false; mult = 1; break;
    case 3: row = sr.s[1]; is_sin = true; mult = 0; break;
    case 4: row = sr.s[1]; is_sin = false; mult = 1; break;
    case 5: row = sr.s[5]; is_sin = true; mult = 2; break;
    case 6: row = sr.s[3]; is_sin = false; mult = 0; break;
    case 7: row = sr.s[6]; is_sin = true; mult = 2; break;
    default: row = sr.s[4]; is_sin = false; mult = 0; break;
  }
}

template <typename R>
__device__ __forceinline__ int field_maxexp(const SlabRows<R>& sr, int f, int K) {
  const R* row;
  bool is_sin;
  int mult;
  field_src(sr, f, row, is_sin, mult);
  return series_maxexp<R>(row, K, is_sin, mult);
}

template <typename R>
__device__ __forceinline__ void syn_pair(cplx<R>* W, const SlabRows<R>& sr, int p, int K, int L,
                                         const cplx<R>* rot, const R* sc) {
  const R *u, *v = nullptr;
  bool us, vs = false;
  int um, vm = 0;
  field_src(sr, 2 * p, u, us, um);
  if (p < 4) field_src(sr, 2 * p + 1, v, vs, vm);
  syn_pack<R>(W, u, us, um, v, vs, vm, K, L, rot, sc[2 * p], p < 4 ? sc[2 * p + 1] : (R)1);
}

template <typename R>
__device__ void project_and_store(cplx<R>* Z, int L, int K, const cplx<R>* rot, R* __restrict__ F,
                                  int b, int i, int nr, const R* inv) {
  // Z[0..L): sine pair (Q, R); Z[L..2L): cosine pair (T, S); inv: unscaling
  R* fq = F + ((size_t)(b * 4 + 0) * nr + i) * K;
  R* fr = F + ((size_t)(b * 4 + 1) * nr + i) * K;
  R* ft = F + ((size_t)(b * 4 + 2) * nr + i) * K;
  R* fs = F + ((size_t)(b * 4 + 3) * nr + i) * K;
  const R invL = (R)1 / (R)L;
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    R c1, c2;
    ana_unpack<R>(Z + L, L, k, rot, c1, c2);
    R a1 = ((R)2 * (c1 * inv[2])) * invL, a2 = ((R)2 * (c2 * inv[3])) * invL;
    if (k == 0) { a1 *= (R)0.5; a2 *= (R)0.5; }
    ft[k] = a1;
    fs[k] = a2;
    R b1 = (R)0, b2 = (R)0;
    if (k > 0) {
      ana_unpack<R>(Z, L, L - k, rot, c1, c2);
      b1 = ((R)2 * (c1 * inv[0])) * invL;
      b2 = ((R)2 * (c2 * inv[1])) * invL;
    }
    fq[k] = b1;
    fr[k] = b2;
  }
}

// Scale the two product pairs (Q,R) at O and (NT,NS) at O+L to unit
// magnitude before the forward FFT; returns the unscaling factors.
template <typename R>
__device__ __forceinline__ void scale_products(cplx<R>* O, int L, int* red, R (&inv)[4]) {
  int mx[4] = {-1000, -1000, -1000, -1000};
  for (int q = threadIdx.x; q < L; q += blockDim.x) {
    mx[0] = max(mx[0], bexp(O[q].x));
    mx[1] = max(mx[1], bexp(O[q].y));
    mx[2] = max(mx[2], bexp(O[L + q].x));
    mx[3] = max(mx[3], bexp(O[L + q].y));
  }
  block_max_exp<4>(mx, red);
  R sc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    sc[k] = pow2_scale_e<R>(mx[k]);
    inv[k] = (R)1 / sc[k];
  }
  for (int q = threadIdx.x; q < L; q += blockDim.x) {
    O[q] = {O[q].x * sc[0], O[q].y * sc[1]};
    O[L + q] = {O[L + q].x * sc[2], O[L + q].y * sc[3]};
  }
  __syncthreads();
}

// Threads of the generic kernel.  It only runs the linearised products of
// grids shorter than 192 points (N_fm below 128; the longer grids have the
// in-place kernels), so four wavefronts cover every loop over the grid, and
// at 1024 threads its scale arrays spilled to scratch (the 128-VGPR limit of
// four wavefronts per SIMD).
constexpr int kNlGenericThreads = 256;

template <typename R, bool DFX>
__global__ __launch_bounds__(kNlGenericThreads) void k_nonlinear(FftShape sh, const cplx<R>* __restrict__ tw_g,
                                                   const cplx<R>* __restrict__ rot_g, int K, int nr,
                                                   const R* __restrict__ slabs,
                                                   const R* __restrict__ slabs2,
                                                   R* __restrict__ F) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int red[16 * 18];
  using C = cplx<R>;
  const int L = sh.L;
  C* A = reinterpret_cast<C*>(smem_raw);
  C* Bf = A + 5 * L;
  // twiddles and half-sample rotations live in LDS next to the data
  C* tw = A + 10 * L;
  C* rot = tw + L;
  for (int k = threadIdx.x; k < L; k += blockDim.x) {
    tw[k] = tw_g[k];
    rot[k] = rot_g[k];
  }
  __syncthreads();
  const int i = blockIdx.x;
  const int b = blockIdx.y;
  // Stage the seven slab rows of X in the second work buffer (idle until the
  // first FFT): one batch of coalesced loads instead of dependent load/use
  // chains in the synthesis loops.
  const SlabRows<R> sg = slab_rows<R>(slabs, b, i, nr, K);
  R* stage = reinterpret_cast<R*>(Bf);
#pragma unroll
  for (int r = 0; r < 7; ++r)
    for (int k = threadIdx.x; k < K; k += blockDim.x) stage[r * K + k] = sg.s[r][k];
  SlabRows<R> sr;
#pragma unroll
  for (int r = 0; r < 7; ++r) sr.s[r] = stage + r * K;
  SlabRows<R> sd = sr;
  if (DFX) sd = slab_rows<R>(slabs2, b, i, nr, K);
  __syncthreads();

  // power-of-two scales of the nine fields of X (and of dv)
  int mx[18];
#pragma unroll
  for (int f = 0; f < 9; ++f) {
    mx[f] = field_maxexp<R>(sr, f, K);
    mx[9 + f] = DFX ? field_maxexp<R>(sd, f, K) : -1000;
  }
  block_max_exp<18>(mx, red);
  R sc[18], is[18];
#pragma unroll
  for (int f = 0; f < 18; ++f) {
    sc[f] = pow2_scale_e<R>(mx[f]);
    is[f] = (R)1 / sc[f];
  }

  // unrolled, here and below: the pair index then selects entries of the
  // scale arrays at compile time and they stay in registers (indexed at run
  // time they were placed in scratch memory)
#pragma unroll
  for (int p = 0; p < 5; ++p) syn_pair<R>(A + p * L, sr, p, K, L, rot, sc);
  __syncthreads();
  C* G = fft_run<R>(A, Bf, 5, sh, tw, true);   // grid fields of X (scaled)
  C* O = (G == A) ? Bf : A;                    // free buffer (5 arrays)
  R inv[4];

  if (!DFX) {
    for (int q = threadIdx.x; q < L; q += blockDim.x) {
      const R sg = (R)grid_sign(q, L);
      const C w0 = G[q], w1 = G[L + q], w2 = G[2 * L + q], w3 = G[3 * L + q], w4 = G[4 * L + q];
      const R JT = w0.x * is[0], Om = sg * w0.y * is[1];
      const R kOm = w1.x * is[2], Dpsi = sg * w1.y * is[3];
      const R kDpsi = w2.x * is[4], kT = sg * w2.y * is[5];
      const R DT = w3.x * is[6], kC = sg * w3.y * is[7];
      const R DC = w4.x * is[8];
      // Matrix_Operators.py:791-793 (Dr applied to JT*omega after projection)
      const R Q = JT * Om;
      const R Rr = kDpsi * Om + Dpsi * kOm;
      const R NT = JT * DT - Dpsi * kT;
      const R NS = JT * DC - Dpsi * kC;
      O[q] = {sg * Q, sg * Rr};
      O[L + q] = {NT, NS};
    }
    __syncthreads();
    scale_products<R>(O, L, red, inv);
    C* Z = fft_run<R>(O, G, 2, sh, tw, false);
    project_and_store<R>(Z, L, K, rot, F, b, i, nr, inv);
    return;
  }

  // Frechet derivative (Matrix_Operators.py:884-887): every product term is
  // (field of X) * (field of dv); the dv fields are synthesised one pair at a
  // time and accumulated.
  C* Zq = O + 2 * L;   // accumulates (Q, R)
  C* Zt = O + 3 * L;   // accumulates (NT, NS)
  for (int q = threadIdx.x; q < L; q += blockDim.x) {
    Zq[q] = {(R)0, (R)0};
    Zt[q] = {(R)0, (R)0};
  }
#pragma unroll
  for (int p = 0; p < 5; ++p) {
    syn_pair<R>(O, sd, p, K, L, rot, sc + 9);
    __syncthreads();
    C* Pg = fft_run<R>(O, O + L, 1, sh, tw, true);
    const R iu = is[9 + 2 * p], iv = p < 4 ? is[9 + 2 * p + 1] : (R)1;
    for (int q = threadIdx.x; q < L; q += blockDim.x) {
      const R sg = (R)grid_sign(q, L);
      const C w0 = G[q], w1 = G[L + q], w2 = G[2 * L + q], w3 = G[3 * L + q], w4 = G[4 * L + q];
      const R JT = w0.x * is[0], Om = sg * w0.y * is[1], kOm = w1.x * is[2],
              Dpsi = sg * w1.y * is[3], kDpsi = w2.x * is[4], kT = sg * w2.y * is[5],
              DT = w3.x * is[6], kC = sg * w3.y * is[7], DC = w4.x * is[8];
      const R du = Pg[q].x * iu, dvv = sg * Pg[q].y * iv;
      C zq = Zq[q], zt = Zt[q];
      if (p == 0) {         // dJT = du, dOm = dvv
        zq.x += JT * dvv + du * Om;
        zq.y += kDpsi * dvv;
        zt.x += du * DT;
        zt.y += du * DC;
      } else if (p == 1) {  // dkOm = du, dDpsi = dvv
        zq.y += Dpsi * du + dvv * kOm;
        zt.x -= dvv * kT;
        zt.y -= dvv * kC;
      } else if (p == 2) {  // dkDpsi = du, dkT = dvv
        zq.y += du * Om;
        zt.x -= Dpsi * dvv;
      } else if (p == 3) {  // dDT = du, dkC = dvv
        zt.x += JT * du;
        zt.y -= Dpsi * dvv;
      } else {              // dDC = du
        zt.y += JT * du;
      }
      if (p == 4) zq = {sg * zq.x, sg * zq.y};
      Zq[q] = zq;
      Zt[q] = zt;
    }
    __syncthreads();
  }
  scale_products<R>(Zq, L, red, inv);
  C* Z = fft_run<R>(Zq, O, 2, sh, tw, false);
  project_and_store<R>(Z, L, K, rot, F, b, i, nr, inv);
}

// ---------------------------------------------------------------------------
// F(X, X) with one wavefront per transform.  A workgroup of five wavefronts
// handles one (radial row i, instance b).  Wavefront p stages the two series
// of field pair p in its own LDS buffers, scales each by a power of two
// (wavefront DPP maximum of exponents), packs them and runs the inverse FFT
// with no workgroup barrier between passes.  After one workgroup barrier,
// wavefront 0 forms the sine products (JT*omega, kDpsi*omega + Dpsi*komega)
// and wavefront 1 the cosine products (T and S), each in a free buffer; each
// runs its forward FFT and projects its pair onto the K modes.
// ---------------------------------------------------------------------------
// Scaled synthesis coefficient a''_k from a staged row c (mode factors
// already applied): cosine index k, or sine index m = L-k (see syn_coef).
template <typename R>
__device__ __forceinline__ R stg_coef(const R* c, int K, int L, int k, bool is_sin) {
  if (!is_sin) return k < K ? (k == 0 ? c[0] : (R)0.5 * c[k]) : (R)0;
  const int m = L - k;
  return (k == 0 || m >= K) ? (R)0 : (R)0.5 * c[m];
}

template <typename R>
__global__ __launch_bounds__(320) void k_nonlinear_w(FftShape sh, const cplx<R>* __restrict__ tw_g,
                                                     const cplx<R>* __restrict__ rot_g, int K,
                                                     int nr, const R* __restrict__ slabs,
                                                     R* __restrict__ F) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ R sinv[10];   // unscaling factors of the nine fields
  __shared__ int sres[5];  // buffer (0 or 1) holding the grid values of pair p
  using C = cplx<R>;
  const int L = sh.L;
  C* tw = reinterpret_cast<C*>(smem_raw);
  C* rot = tw + L;
  C* buf = rot + L;   // [5 pairs][2 buffers][L]
  // Global loads are issued together (unrolled, predicated) before their
  // values are used: a loop that waits for each load costs one memory
  // latency per iteration.
  {
    constexpr int NT = 5;   // covers L <= 1600 with 320 threads
    C t[NT], r[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int kr = threadIdx.x + 320 * j;
      const int k = kr < L ? kr : L - 1;   // clamped: loads unconditional
      t[j] = tw_g[k];
      r[j] = rot_g[k];
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int k = threadIdx.x + 320 * j;
      if (k < L) {
        tw[k] = t[j];
        rot[k] = r[j];
      }
    }
    for (int k = threadIdx.x + 320 * NT; k < L; k += blockDim.x) {
      tw[k] = tw_g[k];
      rot[k] = rot_g[k];
    }
  }
  __syncthreads();
  const int p = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int i = blockIdx.x, b = blockIdx.y;
  const SlabRows<R> sg = slab_rows<R>(slabs, b, i, nr, K);
  {
    C* A = buf + (size_t)p * 2 * L;
    C* Bb = A + L;
    const R *u, *v = nullptr;
    bool us, vs = false;
    int um, vm = 0;
    field_src(sg, 2 * p, u, us, um);
    if (p < 4) field_src(sg, 2 * p + 1, v, vs, vm);
    R* cu = reinterpret_cast<R*>(Bb);   // staged rows; Bb is free until the FFT
    R* cv = cu + K;
    int eu = -1000, ev = -1000;
    auto stage = [&](int k, R a, R c) {
      if (um == 1) a = (R)k * a;
      if (um == 2) a = (R)(-k) * a;
      if (vm == 1) c = (R)k * c;
      if (vm == 2) c = (R)(-k) * c;
      cu[k] = a;
      cv[k] = c;
      if (!(us && k == 0)) eu = max(eu, bexp(a));
      if (!(vs && k == 0)) ev = max(ev, bexp(c));
    };
    constexpr int NJ = 16;   // K <= 1024 in one batch of loads
    R ra[NJ], rc[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int kr = lane + 64 * j;
      const int k = kr < K ? kr : K - 1;   // clamped: loads unconditional
      ra[j] = u[k];
      rc[j] = v ? v[k] : (R)0;
    }
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int k = lane + 64 * j;
      if (k < K) stage(k, ra[j], rc[j]);
    }
    for (int k = lane + 64 * NJ; k < K; k += 64) stage(k, u[k], v ? v[k] : (R)0);
    eu = __ockl_wfred_max_i32(eu);
    ev = __ockl_wfred_max_i32(ev);
    const R su = pow2_scale_e<R>(eu), sv = pow2_scale_e<R>(ev);
    if (lane == 0) {
      sinv[2 * p] = (R)1 / su;
      if (p < 4) sinv[2 * p + 1] = (R)1 / sv;
    }
    wave_lds_sync();
    {
      const R* __restrict__ ccu = cu;
      const R* __restrict__ ccv = cv;
      C* __restrict__ WA = A;
      constexpr int U = 4;
      for (int k0 = lane; k0 < L; k0 += 64 * U) {
        R au[U], av[U], aum[U], avm[U];
        C rr[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int kr = k0 + 64 * u;
          const int k = kr < L ? kr : L - 1;
          au[u] = su * stg_coef(ccu, K, L, k, us);
          av[u] = sv * stg_coef(ccv, K, L, k, vs);
          aum[u] = k ? su * stg_coef(ccu, K, L, L - k, us) : (R)0;
          avm[u] = k ? sv * stg_coef(ccv, K, L, L - k, vs) : (R)0;
          rr[u] = rot[k];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (k0 + 64 * u < L) WA[k0 + 64 * u] = cmul(C{au[u] + avm[u], av[u] - aum[u]}, rr[u]);
      }
    }
    wave_lds_sync();
    C* res = fft_run_w<R>(A, Bb, sh, tw, true, lane);
    if (lane == 0) sres[p] = (res == A) ? 0 : 1;
  }
  __syncthreads();
  if (p >= 2) return;

  C* G[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) G[q] = buf + (size_t)q * 2 * L + (size_t)sres[q] * L;
  R is[9];
#pragma unroll
  for (int f = 0; f < 9; ++f) is[f] = sinv[f];
  C* O = buf + (size_t)p * 2 * L + (size_t)(1 - sres[p]) * L;              // free buffer of pair p
  C* O2 = buf + (size_t)(p + 2) * 2 * L + (size_t)(1 - sres[p + 2]) * L;   // free buffer of pair p+2
  int e0 = -1000, e1 = -1000;
  {
    constexpr int U = 4;
    const C* __restrict__ G0 = G[0];
    const C* __restrict__ G1 = G[1];
    const C* __restrict__ G2 = G[2];
    const C* __restrict__ G3 = G[3];
    const C* __restrict__ G4 = G[4];
    C* __restrict__ Ow = O;
    for (int q0 = lane; q0 < L; q0 += 64 * U) {
      C w0[U], w1[U], w2[U], w3[U], w4[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int qr = q0 + 64 * u;
        const int q = qr < L ? qr : L - 1;
        w0[u] = G0[q];
        w1[u] = G1[q];
        w2[u] = G2[q];
        if (p == 1) {
          w3[u] = G3[q];
          w4[u] = G4[q];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int q = q0 + 64 * u;
        const R gs = (R)grid_sign(q, L);
        const R JT = w0[u].x * is[0], Om = gs * w0[u].y * is[1];
        const R kOm = w1[u].x * is[2], Dpsi = gs * w1[u].y * is[3];
        C o;
        if (p == 0) {
          // Matrix_Operators.py:791 (Dr applied to JT*omega after projection)
          const R kDpsi = w2[u].x * is[4];
          o = {gs * (JT * Om), gs * (kDpsi * Om + Dpsi * kOm)};
        } else {
          // Matrix_Operators.py:792-793
          const R kT = gs * w2[u].y * is[5], DT = w3[u].x * is[6], kC = gs * w3[u].y * is[7],
                  DC = w4[u].x * is[8];
          o = {JT * DT - Dpsi * kT, JT * DC - Dpsi * kC};
        }
        if (q < L) {
          Ow[q] = o;
          e0 = max(e0, bexp(o.x));
          e1 = max(e1, bexp(o.y));
        }
      }
    }
  }
  e0 = __ockl_wfred_max_i32(e0);
  e1 = __ockl_wfred_max_i32(e1);
  const R s0 = pow2_scale_e<R>(e0), s1 = pow2_scale_e<R>(e1);
  {
    constexpr int U = 4;
    C* __restrict__ Ow = O;
    for (int q0 = lane; q0 < L; q0 += 64 * U) {   // the lane's own elements
      C t[U];
#pragma unroll
      for (int u = 0; u < U; ++u) t[u] = Ow[(q0 + 64 * u) < L ? q0 + 64 * u : L - 1];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (q0 + 64 * u < L) Ow[q0 + 64 * u] = {t[u].x * s0, t[u].y * s1};
    }
  }
  wave_lds_sync();
  C* Z = fft_run_w<R>(O, O2, sh, tw, false, lane);
  const R i0 = (R)1 / s0, i1 = (R)1 / s1, invL = (R)1 / (R)L;
  R* f0 = F + ((size_t)(b * 4 + 2 * p) * nr + i) * K;
  R* f1 = F + ((size_t)(b * 4 + 2 * p + 1) * nr + i) * K;
  {
    constexpr int U = 4;
    const C* __restrict__ Zr = Z;
    for (int k0 = lane; k0 < K; k0 += 64 * U) {
      C za[U], zb[U], e[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int kr = k0 + 64 * u;
        const int k = kr < K ? kr : K - 1;
        const int ka = (p == 0) ? L - k : k;   // sine coefficients read index L-k
        const int kb = (L - ka) % L;
        za[u] = Zr[ka % L];
        zb[u] = Zr[kb];
        e[u] = rot[ka % L];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + 64 * u;
        // X = (a + conj b)/2, Y = (a - conj b)/(2i); C = Re(conj(e) X), Re(conj(e) Y)
        const R xr = (R)0.5 * (za[u].x + zb[u].x), xi = (R)0.5 * (za[u].y - zb[u].y);
        const R yr = (R)0.5 * (za[u].y + zb[u].y), yi = (R)-0.5 * (za[u].x - zb[u].x);
        const R c1 = e[u].x * xr + e[u].y * xi;
        const R c2 = e[u].x * yr + e[u].y * yi;
        R a1, a2;
        if (p == 0) {   // sine coefficients (k = 0 is zero)
          a1 = k > 0 ? ((R)2 * (c1 * i0)) * invL : (R)0;
          a2 = k > 0 ? ((R)2 * (c2 * i1)) * invL : (R)0;
        } else {        // cosine coefficients
          a1 = ((R)2 * (c1 * i0)) * invL;
          a2 = ((R)2 * (c2 * i1)) * invL;
          if (k == 0) {
            a1 *= (R)0.5;
            a2 *= (R)0.5;
          }
        }
        if (k < K) {
          f0[k] = a1;
          f1[k] = a2;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// In-place nonlinear kernels, for grids of L = 64*NC points (N_fm = 2L/3 =
// 64*NK modes).  Same work split and arithmetic as k_nonlinear_w, with three
// changes that cut LDS from 12 L to 8 L complex values for F(X, X) (three
// workgroups per CU at L = 768 in float):
//  * each transform runs in place in one buffer (fft_run_ct: the pass
//    schedule fixed at compile time);
//  * the slab rows are not staged in LDS: lane l holds modes l + 64j in
//    registers, the mirrored coefficient at index L - k needed by the
//    Makhoul packing comes from lane (64 - l) mod 64 by a cross-lane read;
//  * the half-sample rotations are read from global memory (issued with the
//    slab loads), not kept in LDS.
// ---------------------------------------------------------------------------

// Loads field pair p of a slab set (one wavefront), scales both series by a
// power of two and writes their packed, rotated spectrum W_k into A.  sinv
// receives the two unscaling factors.
template <typename R, int NC>
__device__ __forceinline__ void ip_pack_pair(const SlabRows<R>& sg, int p, cplx<R>* A,
                                             const cplx<R>* __restrict__ rot_g, R* sinv,
                                             int lane) {
  constexpr int NK = 2 * NC / 3;
  using C = cplx<R>;
  const R *u, *v = nullptr;
  bool us, vs = false;
  int um, vm = 0;
  field_src(sg, 2 * p, u, us, um);
  if (p < 4) field_src(sg, 2 * p + 1, v, vs, vm);
  C rr[NC];
  R du[NK], dv[NK];
#pragma unroll
  for (int j = 0; j < NK; ++j) {
    du[j] = u[lane + 64 * j];
    dv[j] = v ? v[lane + 64 * j] : (R)0;
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) rr[c] = rot_g[lane + 64 * c];
  // mode factors (Derivatives, Matrix_Operators.py:709-729) and exponents
  int eu = -1000, ev = -1000;
#pragma unroll
  for (int j = 0; j < NK; ++j) {
    const int k = lane + 64 * j;
    R a = du[j], c = dv[j];
    if (um == 1) a = (R)k * a;
    if (um == 2) a = (R)(-k) * a;
    if (vm == 1) c = (R)k * c;
    if (vm == 2) c = (R)(-k) * c;
    if (!(us && k == 0)) eu = max(eu, bexp(a));
    if (!(vs && k == 0)) ev = max(ev, bexp(c));
    du[j] = a;
    dv[j] = c;
  }
  eu = __ockl_wfred_max_i32(eu);
  ev = __ockl_wfred_max_i32(ev);
  const R su = pow2_scale_e<R>(eu), sv = pow2_scale_e<R>(ev);
  if (lane == 0) {
    sinv[0] = (R)1 / su;
    if (p < 4) sinv[1] = (R)1 / sv;
  }
  // scaled synthesis coefficients a''_k at mode k = lane + 64 j (see syn_coef;
  // a sine row's index 0 is never used)
#pragma unroll
  for (int j = 0; j < NK; ++j) {
    const bool k0 = (lane == 0 && j == 0);
    du[j] = su * (k0 ? (us ? (R)0 : du[j]) : (R)0.5 * du[j]);
    dv[j] = sv * (k0 ? (vs ? (R)0 : dv[j]) : (R)0.5 * dv[j]);
  }
  // mirrored coefficients at index L - (lane + 64 c): lane (64 - lane) mod 64,
  // register NC-1-c (lane 0: its own register NC-c)
  R mu[NC], mv[NC];
  const int src = (64 - lane) & 63;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    R xu = (R)0, xv = (R)0;
    if (NC - 1 - c < NK) {
      xu = __shfl(du[NC - 1 - c], src);
      xv = __shfl(dv[NC - 1 - c], src);
    }
    if (lane == 0) {
      xu = (NC - c < NK) ? du[NC - c < NK ? NC - c : 0] : (R)0;
      xv = (NC - c < NK) ? dv[NC - c < NK ? NC - c : 0] : (R)0;
    }
    mu[c] = xu;
    mv[c] = xv;
  }
  // packed, rotated spectrum W_k (Makhoul)
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int k = lane + 64 * c;
    const R dcu = c < NK ? du[c < NK ? c : 0] : (R)0;
    const R dcv = c < NK ? dv[c < NK ? c : 0] : (R)0;
    const R au = us ? mu[c] : dcu, av = vs ? mv[c] : dcv;
    const R aum = k ? (us ? dcu : mu[c]) : (R)0;
    const R avm = k ? (vs ? dcv : mv[c]) : (R)0;
    A[k] = cmul(C{au + avm, av - aum}, rr[c]);
  }
}

// Twiddles of the workgroup's transforms: loaded into registers first (in
// flight with the slab loads), stored to LDS after the packing.
template <typename R, int NC, int NTHR>
struct IpTwiddles {
  static constexpr int L = 64 * NC, NT = (L + NTHR - 1) / NTHR;
  cplx<R> t[NT];
  __device__ __forceinline__ void load(const cplx<R>* __restrict__ tw_g) {
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int kr = threadIdx.x + NTHR * j;
      t[j] = tw_g[kr < L ? kr : L - 1];
    }
  }
  __device__ __forceinline__ void store(cplx<R>* tw) const {
#pragma unroll
    for (int j = 0; j < NT; ++j)
      if (threadIdx.x + NTHR * j < L) tw[threadIdx.x + NTHR * j] = t[j];
  }
};

// The second phase for wavefront p in {0, 1}: products o (registers) scaled
// to unit magnitude, written to O, forward FFT in place, projection onto the
// K modes and stores of output rows 2p, 2p+1.
template <typename R, int NC>
__device__ __forceinline__ void ip_forward_project(cplx<R> (&o)[NC], cplx<R>* O, int p,
                                                   const cplx<R>* tw, const FftShape& sh,
                                                   const cplx<R>* __restrict__ rot_g,
                                                   R* __restrict__ F, int b, int i, int nr,
                                                   int lane) {
  constexpr int L = 64 * NC, NK = 2 * NC / 3, K = 64 * NK;
  using C = cplx<R>;
  C er[NK];
#pragma unroll
  for (int j = 0; j < NK; ++j) {
    const int k = lane + 64 * j;
    const int ka = (p == 0) ? L - k : k;   // sine coefficients read index L-k
    er[j] = rot_g[ka < L ? ka : ka - L];
  }
  int e0 = -1000, e1 = -1000;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    e0 = max(e0, bexp(o[c].x));
    e1 = max(e1, bexp(o[c].y));
  }
  e0 = __ockl_wfred_max_i32(e0);
  e1 = __ockl_wfred_max_i32(e1);
  const R s0 = pow2_scale_e<R>(e0), s1 = pow2_scale_e<R>(e1);
#pragma unroll
  for (int c = 0; c < NC; ++c) O[lane + 64 * c] = {o[c].x * s0, o[c].y * s1};
  wave_lds_sync();
  fft_run_ct<R, L>(O, tw, false, lane);
  asm volatile("" : "+v"(lane));   // recompute the projection's indices here
  const R i0 = (R)1 / s0, i1 = (R)1 / s1, invL = (R)1 / (R)L;
  R* f0 = F + ((size_t)(b * 4 + 2 * p) * nr + i) * K;
  R* f1 = F + ((size_t)(b * 4 + 2 * p + 1) * nr + i) * K;
  C za[NK], zb[NK];
#pragma unroll
  for (int j = 0; j < NK; ++j) {
    const int k = lane + 64 * j;
    const int ka = (p == 0) ? L - k : k;
    const int kam = ka < L ? ka : ka - L;
    const int kb = (L - ka) < L ? L - ka : L - ka - L;
    za[j] = O[kam];
    zb[j] = O[kb];
  }
#pragma unroll
  for (int j = 0; j < NK; ++j) {
    const int k = lane + 64 * j;
    // X = (a + conj b)/2, Y = (a - conj b)/(2i); C = Re(conj(e) X), Re(conj(e) Y)
    const R xr = (R)0.5 * (za[j].x + zb[j].x), xi = (R)0.5 * (za[j].y - zb[j].y);
    const R yr = (R)0.5 * (za[j].y + zb[j].y), yi = (R)-0.5 * (za[j].x - zb[j].x);
    const R c1 = er[j].x * xr + er[j].y * xi;
    const R c2 = er[j].x * yr + er[j].y * yi;
    R a1, a2;
    if (p == 0) {   // sine coefficients (k = 0 is zero)
      a1 = k > 0 ? ((R)2 * (c1 * i0)) * invL : (R)0;
      a2 = k > 0 ? ((R)2 * (c2 * i1)) * invL : (R)0;
    } else {        // cosine coefficients
      a1 = ((R)2 * (c1 * i0)) * invL;
      a2 = ((R)2 * (c2 * i1)) * invL;
      if (k == 0) {
        a1 *= (R)0.5;
        a2 *= (R)0.5;
      }
    }
    f0[k] = a1;
    f1[k] = a2;
  }
}

// Waves per SIMD the float kernel's registers must allow: its LDS (five
// pair buffers, 31 KB at L = 768) fits five workgroups of five wavefronts per
// CU, so a SIMD holds up to seven waves (at most 72 VGPRs; left free, the
// compiler took 75, which kept four workgroups per CU).
template <typename R>
constexpr int nl_ip_waves_per_simd() {
  return std::is_same<R, float>::value ? 7 : 1;
}

// F(X, X): five wavefronts, wavefront p synthesises field pair p.
// LDS: five pair buffers of L values each; the twiddles are read from global
// memory (the per-pass table, L1-resident), which leaves room for a fifth
// workgroup per CU (four with them in LDS; measured at L = 768, 64
// instances: 48.0 -> 45.8 us in float, 429 -> 416 us in float64).  Once
// every product is in registers, wavefronts 0 and 1 reuse pair buffers 3
// and 4 for their forward transforms.
template <typename R, int NC>
__global__ __launch_bounds__(320)
__attribute__((amdgpu_waves_per_eu(nl_ip_waves_per_simd<R>()))) void k_nonlinear_ip(
    FftShape sh, const cplx<R>* __restrict__ tw_g, const cplx<R>* __restrict__ rot_g, int nr,
    const R* __restrict__ slabs, R* __restrict__ F) {
  constexpr int L = 64 * NC, NK = 2 * NC / 3, K = 64 * NK;
  static_assert(NC % 3 == 0, "L = 3K/2 with K a multiple of 64");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ R sinv[10];   // unscaling factors of the nine fields
  using C = cplx<R>;
  C* buf = reinterpret_cast<C*>(smem_raw);
  const C* tw = tw_g;
  const int p = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int i = blockIdx.x, b = blockIdx.y;
  const SlabRows<R> sg = slab_rows<R>(slabs, b, i, nr, K);
  C* A = buf + (size_t)p * L;
  ip_pack_pair<R, NC>(sg, p, A, rot_g, sinv + 2 * p, lane);
  __syncthreads();   // the five transforms start together
  fft_run_ct<R, L>(A, tw, true, lane);
  __syncthreads();

  // products of the grid fields (Matrix_Operators.py:791-793)
  const C* G0 = buf;
  const C* G1 = buf + L;
  const C* G2 = buf + 2 * L;
  const C* G3 = buf + 3 * L;
  const C* G4 = buf + 4 * L;
  R is[9];
#pragma unroll
  for (int f = 0; f < 9; ++f) is[f] = sinv[f];
  int ln = lane;
  asm volatile("" : "+v"(ln));   // indices recomputed, not kept from the packing
  C o[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (p >= 2) break;
    const int q = ln + 64 * c;
    const R gs = (R)grid_sign(q, L);
    const C w0 = G0[q], w1 = G1[q], w2 = G2[q];
    const R JT = w0.x * is[0], Om = gs * w0.y * is[1];
    const R kOm = w1.x * is[2], Dpsi = gs * w1.y * is[3];
    if (p == 0) {
      // Matrix_Operators.py:791 (Dr applied to JT*omega after projection)
      const R kDpsi = w2.x * is[4];
      o[c] = {gs * (JT * Om), gs * (kDpsi * Om + Dpsi * kOm)};
    } else {
      // Matrix_Operators.py:792-793
      const C w3 = G3[q], w4 = G4[q];
      const R kT = gs * w2.y * is[5], DT = w3.x * is[6], kC = gs * w3.y * is[7],
              DC = w4.x * is[8];
      o[c] = {JT * DT - Dpsi * kT, JT * DC - Dpsi * kC};
    }
  }
  __syncthreads();   // every read of the pair buffers is done
  if (p >= 2) return;
  ip_forward_project<R, NC>(o, buf + (size_t)(3 + p) * L, p, tw, sh, rot_g, F, b, i, nr,
                                lane);
}

// F(X, dv) + F(dv, X) (NLIN_DFX, Matrix_Operators.py:807-898): ten
// wavefronts, wavefront q < 5 synthesises pair q of X, q >= 5 pair q-5 of dv.
// LDS: twiddles (L), ten pair buffers; wavefronts 0 and 1 reuse pair
// buffers 3 and 4 for their forward transforms once the products are formed.
template <typename R, int NC>
__global__ __launch_bounds__(640) void k_nonlinear_dfx_ip(FftShape sh,
                                                          const cplx<R>* __restrict__ tw_g,
                                                          const cplx<R>* __restrict__ rot_g,
                                                          int nr, const R* __restrict__ slabs,
                                                          const R* __restrict__ slabs2,
                                                          R* __restrict__ F) {
  constexpr int L = 64 * NC, NK = 2 * NC / 3, K = 64 * NK;
  static_assert(NC % 3 == 0, "L = 3K/2 with K a multiple of 64");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ R sinv[20];   // unscaling factors: fields of X, then of dv
  using C = cplx<R>;
  C* tw = reinterpret_cast<C*>(smem_raw);
  C* buf = tw + L;
  const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int i = blockIdx.x, b = blockIdx.y;
  const int p = q < 5 ? q : q - 5;
  const SlabRows<R> sg = slab_rows<R>(q < 5 ? slabs : slabs2, b, i, nr, K);
  IpTwiddles<R, NC, 640> twr;
  twr.load(tw_g);
  C* A = buf + (size_t)q * L;
  ip_pack_pair<R, NC>(sg, p, A, rot_g, sinv + (q < 5 ? 0 : 10) + 2 * p, lane);
  twr.store(tw);
  __syncthreads();   // twiddles in LDS
  fft_run_ct<R, L>(A, tw, true, lane);
  __syncthreads();

  // Frechet derivative (Matrix_Operators.py:884-887): every term is (field of
  // X) * (field of dv), accumulated in the order of the pairs of dv
  R is[9], id[9];
#pragma unroll
  for (int f = 0; f < 9; ++f) {
    is[f] = sinv[f];
    id[f] = sinv[10 + f];
  }
  int ln = lane;
  asm volatile("" : "+v"(ln));   // indices recomputed, not kept from the packing
  C o[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (q >= 2) break;
    const int qq = ln + 64 * c;
    const R gs = (R)grid_sign(qq, L);
    const C w0 = buf[qq], w1 = buf[L + qq], w2 = buf[2 * L + qq];
    const C v0 = buf[5 * L + qq], v1 = buf[6 * L + qq], v2 = buf[7 * L + qq];
    const R JT = w0.x * is[0], Om = gs * w0.y * is[1];
    const R kOm = w1.x * is[2], Dpsi = gs * w1.y * is[3];
    const R dJT = v0.x * id[0], dOm = gs * v0.y * id[1];
    const R dkOm = v1.x * id[2], dDpsi = gs * v1.y * id[3];
    const R dkDpsi = v2.x * id[4];
    if (q == 0) {
      const R kDpsi = w2.x * is[4];
      R zx = JT * dOm + dJT * Om;
      R zy = kDpsi * dOm;
      zy += Dpsi * dkOm + dDpsi * kOm;
      zy += dkDpsi * Om;
      o[c] = {gs * zx, gs * zy};
    } else {
      const C w3 = buf[3 * L + qq], w4 = buf[4 * L + qq];
      const C v3 = buf[8 * L + qq], v4 = buf[9 * L + qq];
      const R kT = gs * w2.y * is[5], DT = w3.x * is[6], kC = gs * w3.y * is[7],
              DC = w4.x * is[8];
      const R dkT = gs * v2.y * id[5], dDT = v3.x * id[6], dkC = gs * v3.y * id[7],
              dDC = v4.x * id[8];
      R tx = dJT * DT, ty = dJT * DC;
      tx -= dDpsi * kT;
      ty -= dDpsi * kC;
      tx -= Dpsi * dkT;
      tx += JT * dDT;
      ty -= Dpsi * dkC;
      ty += JT * dDC;
      o[c] = {tx, ty};
    }
  }
  __syncthreads();   // every read of the pair buffers is done
  if (q >= 2) return;
  ip_forward_project<R, NC>(o, buf + (size_t)(3 + q) * L, q, tw, sh, rot_g, F, b, i, nr,
                                lane);
}

// Columns per lane NC of the in-place kernel for this plan, 0 when its
// grid does not fit it (k_nonlinear_w is used then).
static int nl_inplace_nc(const Plan& p) {
  const int L = p.tabM.shape.L;
  if (L % 64 != 0 || 3 * p.K != 2 * L) return 0;
  const int nc = L / 64;
  if (nc != 3 && nc != 6 && nc != 12) return 0;
  return ip_shape_ok(p.tabM.shape) ? nc : 0;
}

// Raises a kernel's dynamic-LDS limit to `bytes` once.  `current` is the
// limit already set for that kernel; calls may come from several host
// threads (the attribute is idempotent, the record is atomic).
int ensure_dynamic_lds(const void* fn, size_t bytes, std::atomic<size_t>& current) {
  if (bytes <= 64 * 1024 || bytes <= current.load(std::memory_order_acquire)) return SDDC_OK;
  std::lock_guard<std::recursive_mutex> lk(device_state_mutex());
  if (bytes <= current.load(std::memory_order_acquire)) return SDDC_OK;
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();   // do not leave the error for the next launch
    return fail(SDDC_ERR_UNSUPPORTED, "dynamic LDS request above the device limit");
  }
  size_t prev = current.load(std::memory_order_relaxed);
  while (prev < bytes && !current.compare_exchange_weak(prev, bytes, std::memory_order_acq_rel)) {
  }
  return SDDC_OK;
}

// Dynamic LDS of the in-place kernels: twiddles (L) and the pair
// buffers (five for F(X, X), ten for the Frechet derivative).
template <typename R>
static size_t nl_ip_lds(int L, bool dfx = false) {
  // F(X, X): five pair buffers; F(X, dv) + F(dv, X): ten, and the twiddles
  return (dfx ? 11 : 5) * (size_t)L * sizeof(cplx<R>);
}

int nonlinear_lds_bytes(const Plan& p, int precision, int /*dfx*/) {
  const size_t c = precision == SDDC_MIXED ? sizeof(float2) : sizeof(double2);
  // in place: 5 pair buffers (the products reuse two); else 2 x 5 working
  // arrays + twiddles + rotations
  if (nl_inplace_nc(p))
    return (int)(precision == SDDC_MIXED ? nl_ip_lds<float>(p.M) : nl_ip_lds<double>(p.M));
  return (int)(12 * (size_t)p.M * c);
}

template <typename R, int NC>
static int launch_nl_ip(const Plan& p, int B, const void* slabs, void* F, hipStream_t s,
                        const cplx<R>* tw, const cplx<R>* rot) {
  const size_t lds = nl_ip_lds<R>(64 * NC);
  static std::atomic<size_t> cur{0};
  if (int rc = ensure_dynamic_lds((const void*)k_nonlinear_ip<R, NC>, lds, cur)) return rc;
  hipLaunchKernelGGL((k_nonlinear_ip<R, NC>), dim3(p.nr, B), dim3(320), lds, s, p.tabM.shape, tw,
                     rot, p.nr, (const R*)slabs, (R*)F);
  return hip_check(hipGetLastError(), "k_nonlinear_ip");
}

template <typename R, int NC>
static int launch_nl_dfx_ip(const Plan& p, int B, const void* slabs, const void* slabs2, void* F,
                            hipStream_t s, const cplx<R>* tw, const cplx<R>* rot) {
  const size_t lds = nl_ip_lds<R>(64 * NC, true);
  static std::atomic<size_t> cur{0};
  if (int rc = ensure_dynamic_lds((const void*)k_nonlinear_dfx_ip<R, NC>, lds, cur)) return rc;
  hipLaunchKernelGGL((k_nonlinear_dfx_ip<R, NC>), dim3(p.nr, B), dim3(640), lds, s, p.tabM.shape,
                     tw, rot, p.nr, (const R*)slabs, (const R*)slabs2, (R*)F);
  return hip_check(hipGetLastError(), "k_nonlinear_dfx_ip");
}

template <typename R, bool DFX>
static int launch_nl_t(const Plan& p, int B, const void* slabs, const void* slabs2, void* F,
                       hipStream_t s, const cplx<R>* tw, const cplx<R>* rot) {
  const size_t lds = 12 * (size_t)p.M * sizeof(cplx<R>);
  // the in-place kernels take the per-pass twiddle layout
  const cplx<R>* twp = std::is_same<R, float>::value
                           ? reinterpret_cast<const cplx<R>*>(p.tabM.twp32)
                           : reinterpret_cast<const cplx<R>*>(p.tabM.twp64);
  if (!DFX) {
    switch (nl_inplace_nc(p)) {
      case 3: return launch_nl_ip<R, 3>(p, B, slabs, F, s, twp, rot);
      case 6: return launch_nl_ip<R, 6>(p, B, slabs, F, s, twp, rot);
      case 12: return launch_nl_ip<R, 12>(p, B, slabs, F, s, twp, rot);
      default: break;
    }
    if (lds + 64 > 160 * 1024)
      return fail(SDDC_ERR_UNSUPPORTED, "N_fm too large for the LDS nonlinear kernel");
    static std::atomic<size_t> curw{0};
    if (int rc = ensure_dynamic_lds((const void*)k_nonlinear_w<R>, lds, curw)) return rc;
    hipLaunchKernelGGL((k_nonlinear_w<R>), dim3(p.nr, B), dim3(320), lds, s, p.tabM.shape, tw,
                       rot, p.K, p.nr, (const R*)slabs, (R*)F);
    return hip_check(hipGetLastError(), "k_nonlinear_w");
  }
  switch ((nl_ip_lds<R>(p.M, true) <= 160 * 1024) ? nl_inplace_nc(p) : 0) {
    case 3: return launch_nl_dfx_ip<R, 3>(p, B, slabs, slabs2, F, s, twp, rot);
    case 6: return launch_nl_dfx_ip<R, 6>(p, B, slabs, slabs2, F, s, twp, rot);
    case 12: return launch_nl_dfx_ip<R, 12>(p, B, slabs, slabs2, F, s, twp, rot);
    default: break;
  }
  if (lds + 16 * 18 * sizeof(int) > 160 * 1024)
    return fail(SDDC_ERR_UNSUPPORTED, "N_fm too large for the LDS nonlinear kernel");
  static std::atomic<size_t> cur{0};
  if (int rc = ensure_dynamic_lds((const void*)k_nonlinear<R, DFX>, lds, cur)) return rc;
  hipLaunchKernelGGL((k_nonlinear<R, DFX>), dim3(p.nr, B), dim3(kNlGenericThreads), lds, s,
                     p.tabM.shape, tw,
                     rot, p.K, p.nr, (const R*)slabs, (const R*)slabs2, (R*)F);
  return hip_check(hipGetLastError(), "k_nonlinear");
}

int launch_nonlinear(const Plan& p, int B, const void* slabs, const void* slabs2, void* F,
                     hipStream_t s) {
  if (B <= 0) return SDDC_OK;
  const bool dfx = slabs2 != nullptr;
  if (p.precision == SDDC_MIXED) {
    auto tw = reinterpret_cast<const cplx<float>*>(p.tabM.tw32);
    auto rot = reinterpret_cast<const cplx<float>*>(p.tabM.rot32);
    return dfx ? launch_nl_t<float, true>(p, B, slabs, slabs2, F, s, tw, rot)
               : launch_nl_t<float, false>(p, B, slabs, slabs2, F, s, tw, rot);
  }
  auto tw = reinterpret_cast<const cplx<double>*>(p.tabM.tw64);
  auto rot = reinterpret_cast<const cplx<double>*>(p.tabM.rot64);
  return dfx ? launch_nl_t<double, true>(p, B, slabs, slabs2, F, s, tw, rot)
             : launch_nl_t<double, false>(p, B, slabs, slabs2, F, s, tw, rot);
}

// ---------------------------------------------------------------------------
// Kinetic energy, one workgroup per (radial row i, instance b): synthesise
// J_theta(psi)/r (cosine) and Dr psi (sine) on grid(3K), and integrate
// (J^2 + D^2) sin(theta) over theta with the trapezoid rule
// (Main.py:107-130).  The radial trapezoid is done in k_diag_final.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_ke_rows(FftShape sh, const double2* tw_,
                                                 const double2* rot_, int K, int nr,
                                                 const double* __restrict__ slabs,
                                                 const double* __restrict__ vecs, int NRP,
                                                 double* __restrict__ rowval) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using C = cplx<double>;
  __shared__ double red[256];
  __shared__ int red2[16 * 2];
  const int L = sh.L;
  C* A = reinterpret_cast<C*>(smem_raw);
  C* Bf = A + L;
  const C* tw = reinterpret_cast<const C*>(tw_);
  const C* rot = reinterpret_cast<const C*>(rot_);
  const int i = blockIdx.x, b = blockIdx.y;
  const double* jt = slabs + ((size_t)(b * 7 + 0) * nr + i) * K;
  const double* dp = slabs + ((size_t)(b * 7 + 1) * nr + i) * K;
  // IR@JPSI = (1/R_i) * JPSI (Main.py:102,112)
  const double ir = vecs[V_IR * NRP + i];
  int mx[2] = {-1000, -1000};
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    mx[0] = max(mx[0], bexp(ir * jt[k]));
    if (k > 0) mx[1] = max(mx[1], bexp(dp[k]));
  }
  block_max_exp<2>(mx, red2);
  const double su = pow2_scale_e<double>(mx[0]), sv = pow2_scale_e<double>(mx[1]);
  for (int k = threadIdx.x; k < L; k += blockDim.x) {
    const double au = k < K ? (k == 0 ? ir * jt[0] : 0.5 * (ir * jt[k])) : 0.0;
    const double aum = (k > 0 && L - k < K) ? 0.5 * (ir * jt[L - k]) : 0.0;
    const int m = L - k;
    const double av = (k > 0 && m < K) ? 0.5 * dp[m] : 0.0;
    const double avm = (k > 0 && k < K) ? 0.5 * dp[k] : 0.0;
    C w = {su * au + sv * avm, sv * av - su * aum};
    A[k] = cmul(w, rot[k]);
  }
  __syncthreads();
  C* res = fft_run<double>(A, Bf, 1, sh, tw, true);
  const double iu = 1.0 / su, iv = 1.0 / sv;
  const double dth = 3.14159265358979323846 / L;
  double acc = 0.0;
  for (int q = threadIdx.x; q < L; q += blockDim.x) {
    const int n = grid_index(q, L);
    const double th = (3.14159265358979323846 / L) * (2.0 * n + 1.0) / 2;
    const double J = res[q].x * iu;
    const double D = res[q].y * iv;  // the (-1)^n sign vanishes in the square
    const double w = (n == 0 || n == L - 1) ? 0.5 * dth : dth;
    acc += w * ((J * J + D * D) * sin(th));
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) rowval[(size_t)b * nr + i] = red[0];
}

int launch_ke_rows(const Plan& p, int B, const void* slabs64, double* rowval, hipStream_t s) {
  const size_t lds = 2 * (size_t)p.tabKE.L * sizeof(double2);
  if (lds > 160 * 1024) return fail(SDDC_ERR_UNSUPPORTED, "N_fm too large for the LDS energy kernel");
  static std::atomic<size_t> cur{0};
  if (int rc = ensure_dynamic_lds((const void*)k_ke_rows, lds, cur)) return rc;
  hipLaunchKernelGGL(k_ke_rows, dim3(p.nr, B), dim3(256), lds, s, p.tabKE.shape, p.tabKE.tw64,
                     p.tabKE.rot64, p.K, p.nr, (const double*)slabs64, p.vecs, p.NRP, rowval);
  return hip_check(hipGetLastError(), "k_ke_rows");
}

}  // namespace sddc
