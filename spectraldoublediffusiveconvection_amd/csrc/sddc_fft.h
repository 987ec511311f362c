// Batched real-to-real cosine/sine transforms on the midpoint grid
// theta_n = pi(2n+1)/(2L), computed in LDS by a length-L complex FFT.
//
// Synthesis (DCT-III, the reference IDCT/IDST, Transforms.py:87-129): two real
// series u, v share one complex FFT.  With a''_k the reference's scaled input
// (a''_0 = a_0, a''_k = a_k/2) and a''_L = 0,
//   W_k = [(a''u_k + a''v_{L-k}) + i (a''v_k - a''u_{L-k})] e^{i pi k/(2L)},
//   w   = inverse FFT_L(W)          (w_q = sum_k W_k e^{+2 pi i k q/L}),
// and Re w_q, Im w_q are the two syntheses at grid point n(q), where
// n(q) = 2q for 2q < L and 2(L-1-q)+1 otherwise (Makhoul's even/odd reorder).
// A sine series b (index m <-> sin m theta) is the cosine series with
// coefficients a_k = b_{L-k} times (-1)^n, because
// sin(m theta_n) = (-1)^n cos((L-m) theta_n).
//
// Analysis (DCT-II, the reference DCT/DST, Transforms.py:73-114): two real grid
// functions x, y in the same reordered layout form z_q = x_{n(q)} + i y_{n(q)},
// Z = forward FFT_L(z), X_k = (Z_k + conj Z_{L-k})/2, Y_k = (Z_k - conj Z_{L-k})/(2i),
// C_k = Re(e^{-i pi k/(2L)} X_k) = sum_n x_n cos(k theta_n).  Sine coefficients
// come from the cosine analysis of (-1)^n x read at index L-m.
//
// Grid data therefore never leaves the reordered layout: the nonlinear
// products are pointwise, so they are formed directly on w.
#pragma once
#include "sddc_internal.h"

namespace sddc {

template <typename R>
__device__ __forceinline__ cplx<R> tw_get(const cplx<R>* tw, int t, bool inv) {
  cplx<R> w = tw[t];
  if (inv) w.y = -w.y;
  return w;
}

// small DFTs, forward sign exp(-2 pi i/r); inv selects exp(+2 pi i/r)
template <typename R>
__device__ __forceinline__ void dft2(cplx<R>* v) {
  cplx<R> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename R>
__device__ __forceinline__ void dft3(cplx<R>* v, bool inv) {
  const R h = (R)0.86602540378443864676;  // sqrt(3)/2
  cplx<R> t = cadd(v[1], v[2]);
  cplx<R> d = csub(v[1], v[2]);
  cplx<R> m = {v[0].x - (R)0.5 * t.x, v[0].y - (R)0.5 * t.y};
  // forward: y1 = m - i h d, y2 = m + i h d
  R s = inv ? -h : h;
  cplx<R> y0 = cadd(v[0], t);
  cplx<R> y1 = {m.x + s * d.y, m.y - s * d.x};
  cplx<R> y2 = {m.x - s * d.y, m.y + s * d.x};
  v[0] = y0;
  v[1] = y1;
  v[2] = y2;
}

template <typename R>
__device__ __forceinline__ void dft4(cplx<R>* v, bool inv) {
  cplx<R> t0 = cadd(v[0], v[2]);
  cplx<R> t1 = csub(v[0], v[2]);
  cplx<R> t2 = cadd(v[1], v[3]);
  cplx<R> t3 = csub(v[1], v[3]);
  // forward: y1 = t1 - i t3, y3 = t1 + i t3
  cplx<R> mi = inv ? cplx<R>{-t3.y, t3.x} : cplx<R>{t3.y, -t3.x};  // (+/- i) t3
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, mi);
  v[3] = csub(t1, mi);
}

template <typename R>
__device__ __forceinline__ void dft5(cplx<R>* v, bool inv) {
  const R c1 = (R)0.30901699437494742410;   // cos(2 pi/5)
  const R c2 = (R)-0.80901699437494742410;  // cos(4 pi/5)
  const R s1 = (R)0.95105651629515357212;   // sin(2 pi/5)
  const R s2 = (R)0.58778525229247312917;   // sin(4 pi/5)
  R sg = inv ? (R)-1 : (R)1;
  cplx<R> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
  cplx<R> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
  cplx<R> y0 = {v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y};
  cplx<R> r1 = {v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y};
  cplx<R> r2 = {v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y};
  // forward: y_k = r_k - i sg (s terms)
  cplx<R> q1 = {s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
  cplx<R> q2 = {s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
  v[0] = y0;
  v[1] = {r1.x + sg * q1.y, r1.y - sg * q1.x};
  v[4] = {r1.x - sg * q1.y, r1.y + sg * q1.x};
  v[2] = {r2.x + sg * q2.y, r2.y - sg * q2.x};
  v[3] = {r2.x - sg * q2.y, r2.y + sg * q2.x};
}

// Quotient w/d for 0 <= w < 2^20 from a float reciprocal of d, corrected to
// exact (integer division proper costs dozens of instructions per use).
__device__ __forceinline__ int qdiv(int w, int d, float rd) {
  int q = (int)((float)w * rd);
  const int r = w - q * d;
  q += (r >= d) - (r < 0);
  return q;
}

// One Stockham autosort pass of radix RADIX over nfft arrays of length L,
// out of place.  Ns is the product of the radices already applied.
template <typename R, int RADIX>
__device__ void fft_pass(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out, int L,
                         int nfft, int Ns, const cplx<R>* __restrict__ tw, bool inv) {
  const int nb = L / RADIX;
  const int st = L / (Ns * RADIX);
  const float rnb = 1.0f / (float)nb, rns = 1.0f / (float)Ns;
  for (int w = threadIdx.x; w < nfft * nb; w += blockDim.x) {
    const int f = qdiv(w, nb, rnb);
    const int j = w - f * nb;
    const cplx<R>* src = in + f * L;
    cplx<R>* dst = out + f * L;
    const int jq = qdiv(j, Ns, rns);
    const int k = j - jq * Ns;
    cplx<R> v[RADIX];
    v[0] = src[j];
#pragma unroll
    for (int s = 1; s < RADIX; ++s) v[s] = cmul(src[j + s * nb], tw_get(tw, s * k * st, inv));
    if (RADIX == 2) dft2(v);
    if (RADIX == 3) dft3(v, inv);
    if (RADIX == 4) dft4(v, inv);
    if (RADIX == 5) dft5(v, inv);
    const int base = jq * Ns * RADIX + k;
#pragma unroll
    for (int s = 0; s < RADIX; ++s) dst[base + s * Ns] = v[s];
  }
}

// Any radix r (used for prime factors above 5): direct r-point DFT.
template <typename R>
__device__ void fft_pass_any(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out, int L,
                             int nfft, int Ns, int r, const cplx<R>* __restrict__ tw, bool inv) {
  const int nb = L / r;
  const int st = L / (Ns * r);
  for (int w = threadIdx.x; w < nfft * nb * r; w += blockDim.x) {
    const int s = w % r;              // output index of the small DFT
    const int wb = w / r;
    const int f = wb / nb;
    const int j = wb - f * nb;
    const cplx<R>* src = in + f * L;
    cplx<R>* dst = out + f * L;
    const int k = j % Ns;
    cplx<R> acc = {(R)0, (R)0};
    for (int t = 0; t < r; ++t) {
      cplx<R> x = src[j + t * nb];
      if (t) x = cmul(x, tw_get(tw, t * k * st, inv));
      acc = cadd(acc, cmul(x, tw_get(tw, ((s * t) % r) * nb, inv)));
    }
    const int base = (j / Ns) * Ns * r + k;
    dst[base + s * Ns] = acc;
  }
}

// Runs all passes; a and b are the two LDS buffers (nfft*L each).  Returns the
// buffer holding the result.  Ends with a barrier.
template <typename R>
__device__ cplx<R>* fft_run(cplx<R>* a, cplx<R>* b, int nfft, const FftShape& sh,
                            const cplx<R>* __restrict__ tw, bool inv) {
  int Ns = 1;
  for (int p = 0; p < sh.npass; ++p) {
    const int r = sh.radix[p];
    if (r == 4)
      fft_pass<R, 4>(a, b, sh.L, nfft, Ns, tw, inv);
    else if (r == 2)
      fft_pass<R, 2>(a, b, sh.L, nfft, Ns, tw, inv);
    else if (r == 3)
      fft_pass<R, 3>(a, b, sh.L, nfft, Ns, tw, inv);
    else if (r == 5)
      fft_pass<R, 5>(a, b, sh.L, nfft, Ns, tw, inv);
    else
      fft_pass_any<R>(a, b, sh.L, nfft, Ns, r, tw, inv);
    __syncthreads();
    cplx<R>* t = a;
    a = b;
    b = t;
    Ns *= r;
  }
  return a;
}

// Sign (-1)^n of the grid point held at reordered position q.
__device__ __forceinline__ int grid_sign(int q, int L) { return (2 * q < L) ? 1 : -1; }
__device__ __forceinline__ int grid_index(int q, int L) {
  return (2 * q < L) ? 2 * q : 2 * (L - 1 - q) + 1;
}

// Scaled synthesis coefficient a''_k of a series row (Transforms.py:16-26,
// 41-54).  mult: 0 -> c, 1 -> m*c, 2 -> -m*c (the theta-derivative factors of
// Derivatives, Matrix_Operators.py:709,714,722,729).
template <typename R>
__device__ __forceinline__ R syn_coef(const R* __restrict__ row, int K, int L, int k,
                                      bool is_sin, int mult) {
  if (row == nullptr) return (R)0;
  if (!is_sin) {
    if (k >= K) return (R)0;
    R c = row[k];
    if (mult == 1) c = (R)k * c;
    if (mult == 2) c = (R)(-k) * c;
    return k == 0 ? c : (R)0.5 * c;
  }
  const int m = L - k;                 // sin(m theta), m in [1, K-1]
  if (k == 0 || m >= K) return (R)0;
  R c = row[m];
  if (mult == 1) c = (R)m * c;
  if (mult == 2) c = (R)(-m) * c;
  return (R)0.5 * c;
}

// Pairing two real transforms in one complex FFT leaves each with rounding
// error relative to the larger of the two.  Each series is therefore scaled
// by a power of two (exact) to unit magnitude before packing, and unscaled
// afterwards.
extern "C" __device__ int __ockl_wfred_max_i32(int);
// Binary exponent e with |x| < 2^e (frexp); zero maps to a large negative value.
template <typename R>
__device__ __forceinline__ int bexp(R x) {
  int e;
  (void)frexp(x, &e);
  return x == (R)0 ? -1000 : e;
}

// Power of two 2^-e bringing magnitudes below 2^e into [0.5, 1).
template <typename R>
__device__ __forceinline__ R pow2_scale_e(int e) {
  if (e < -900 || e > 900) return (R)1;
  return ldexp((R)1, -e);
}

// Block-wide maximum of NV exponents: DPP wavefront reduction, then one LDS
// word per wavefront.  red needs (blockDim/64)*NV ints.  Ends with a barrier.
template <int NV>
__device__ __forceinline__ void block_max_exp(int (&e)[NV], int* red) {
#pragma unroll
  for (int k = 0; k < NV; ++k) e[k] = __ockl_wfred_max_i32(e[k]);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) red[w * NV + k] = e[k];
  __syncthreads();
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    int m = -1000;
    for (int q = 0; q < nw; ++q) m = max(m, red[q * NV + k]);
    e[k] = m;
  }
  __syncthreads();
}

// Largest exponent of |a''_k| over a series row held by this thread's part
// of the row (see syn_coef); combine with block_max_exp.
template <typename R>
__device__ __forceinline__ int series_maxexp(const R* __restrict__ row, int K, bool is_sin,
                                             int mult) {
  int m = -1000;
  if (row == nullptr) return m;
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    if (is_sin && k == 0) continue;
    R c = row[k];
    if (mult) c = (R)k * c;
    m = max(m, bexp(c));
  }
  return m;
}

// Pre-rotated synthesis spectrum of the pair (u*su, v*sv) into W[0..L).
template <typename R>
__device__ __forceinline__ void syn_pack(cplx<R>* __restrict__ W, const R* u, bool u_sin,
                                         int u_mult, const R* v, bool v_sin, int v_mult, int K,
                                         int L, const cplx<R>* __restrict__ rot, R su = (R)1,
                                         R sv = (R)1) {
  for (int k = threadIdx.x; k < L; k += blockDim.x) {
    const R au = su * syn_coef(u, K, L, k, u_sin, u_mult);
    const R av = sv * syn_coef(v, K, L, k, v_sin, v_mult);
    const R aum = k ? su * syn_coef(u, K, L, L - k, u_sin, u_mult) : (R)0;
    const R avm = k ? sv * syn_coef(v, K, L, L - k, v_sin, v_mult) : (R)0;
    cplx<R> w = {au + avm, av - aum};
    W[k] = cmul(w, rot[k]);
  }
}

// Cosine analysis values C_k of the two inputs packed in Z (after the forward
// FFT), at index k.
template <typename R>
__device__ __forceinline__ void ana_unpack(const cplx<R>* __restrict__ Z, int L, int k,
                                           const cplx<R>* __restrict__ rot, R& c1, R& c2) {
  const cplx<R> a = Z[k];
  const cplx<R> b = Z[(L - k) % L];
  // X = (a + conj b)/2 ; Y = (a - conj b)/(2i)
  const R xr = (R)0.5 * (a.x + b.x), xi = (R)0.5 * (a.y - b.y);
  const R dr = a.x - b.x, di = a.y + b.y;
  const R yr = (R)0.5 * di, yi = (R)-0.5 * dr;
  const cplx<R> e = rot[k];  // e^{+i pi k/(2L)}; we need Re(conj(e) * .)
  c1 = e.x * xr + e.y * xi;
  c2 = e.x * yr + e.y * yi;
}

// ---------------------------------------------------------------------------
// Wavefront-local FFT: one wavefront runs every pass of one transform in its
// own LDS buffers.  LDS instructions of a wavefront execute in order, so a
// pass only needs its own LDS writes complete (lgkmcnt(0)) before the next
// pass reads them; no workgroup barrier.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_lds_sync() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0); vmcnt, expcnt not waited on
  __builtin_amdgcn_wave_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Iterations are processed in chunks of U: every LDS read of a chunk (data
// and twiddles, indices clamped in range) is issued before any is used, so
// one wavefront keeps several reads in flight; only stores are predicated.
template <typename R, int RADIX>
__device__ void fft_pass_w(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out, int L,
                           int Ns, const cplx<R>* __restrict__ tw, bool inv, int lane) {
  constexpr int U = RADIX >= 5 ? 2 : (RADIX == 4 ? 3 : 4);
  const int nb = L / RADIX;
  const int st = L / (Ns * RADIX);
  const float rns = 1.0f / (float)Ns;
  for (int j0 = lane; j0 < nb; j0 += 64 * U) {
    cplx<R> v[U][RADIX], w[U][RADIX];
    int base[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int jr = j0 + 64 * u;
      const int j = jr < nb ? jr : nb - 1;
      const int jq = qdiv(j, Ns, rns);
      const int k = j - jq * Ns;
      base[u] = jq * Ns * RADIX + k;
#pragma unroll
      for (int s = 0; s < RADIX; ++s) v[u][s] = in[j + s * nb];
#pragma unroll
      for (int s = 1; s < RADIX; ++s) w[u][s] = tw_get(tw, s * k * st, inv);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int s = 1; s < RADIX; ++s) v[u][s] = cmul(v[u][s], w[u][s]);
      if (RADIX == 2) dft2(v[u]);
      if (RADIX == 3) dft3(v[u], inv);
      if (RADIX == 4) dft4(v[u], inv);
      if (RADIX == 5) dft5(v[u], inv);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (j0 + 64 * u < nb)
#pragma unroll
        for (int s = 0; s < RADIX; ++s) out[base[u] + s * Ns] = v[u][s];
  }
}

template <typename R>
__device__ void fft_pass_any_w(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out, int L,
                               int Ns, int r, const cplx<R>* __restrict__ tw, bool inv, int lane) {
  const int nb = L / r;
  const int st = L / (Ns * r);
  for (int w = lane; w < nb * r; w += 64) {
    const int s = w % r;
    const int j = w / r;
    const int k = j % Ns;
    cplx<R> acc = {(R)0, (R)0};
    for (int t = 0; t < r; ++t) {
      cplx<R> x = in[j + t * nb];
      if (t) x = cmul(x, tw_get(tw, t * k * st, inv));
      acc = cadd(acc, cmul(x, tw_get(tw, ((s * t) % r) * nb, inv)));
    }
    out[(j / Ns) * Ns * r + k + s * Ns] = acc;
  }
}

// All passes of one transform by one wavefront; returns the buffer holding
// the result (a or b).
template <typename R>
__device__ cplx<R>* fft_run_w(cplx<R>* a, cplx<R>* b, const FftShape& sh,
                              const cplx<R>* __restrict__ tw, bool inv, int lane) {
  int Ns = 1;
  for (int p = 0; p < sh.npass; ++p) {
    const int r = sh.radix[p];
    if (r == 4)
      fft_pass_w<R, 4>(a, b, sh.L, Ns, tw, inv, lane);
    else if (r == 2)
      fft_pass_w<R, 2>(a, b, sh.L, Ns, tw, inv, lane);
    else if (r == 3)
      fft_pass_w<R, 3>(a, b, sh.L, Ns, tw, inv, lane);
    else if (r == 5)
      fft_pass_w<R, 5>(a, b, sh.L, Ns, tw, inv, lane);
    else
      fft_pass_any_w<R>(a, b, sh.L, Ns, r, tw, inv, lane);
    wave_lds_sync();
    cplx<R>* t = a;
    a = b;
    b = t;
    Ns *= r;
  }
  return a;
}

// ---------------------------------------------------------------------------
// In-place wavefront FFT.  Each pass reads its whole array into registers
// (one chunk of kIpChunk<RADIX> butterflies per lane covers all L/RADIX) before
// it writes, so input and output share one buffer: half the LDS of the
// ping-pong form.  LDS instructions of a wavefront issue in program order
// and the buffer is not restrict-qualified, so the compiler keeps every read
// of a pass ahead of its writes.  ip_shape_ok checks the chunk condition.
// ---------------------------------------------------------------------------
template <int RADIX>
struct IpChunk {
  static constexpr int U = RADIX >= 5 ? 2 : (RADIX == 4 ? 3 : 4);
};

inline bool ip_shape_ok(const FftShape& sh) {
  if (sh.npass <= 0) return false;
  for (int q = 0; q < sh.npass; ++q) {
    const int r = sh.radix[q];
    int U = 0;
    if (r == 2) U = IpChunk<2>::U;
    if (r == 3) U = IpChunk<3>::U;
    if (r == 4) U = IpChunk<4>::U;
    if (r == 5) U = IpChunk<5>::U;
    if (U == 0 || sh.L / r > 64 * U) return false;
  }
  return true;
}

template <typename R, int RADIX>
__device__ __forceinline__ void fft_pass_ip(cplx<R>* buf, int L, int Ns,
                                            const cplx<R>* __restrict__ tw, bool inv, int lane) {
  constexpr int U = IpChunk<RADIX>::U;
  const int nb = L / RADIX;
  const int st = L / (Ns * RADIX);
  const float rns = 1.0f / (float)Ns;
  cplx<R> v[U][RADIX], w[U][RADIX];
  int base[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int jr = lane + 64 * u;
    const int j = jr < nb ? jr : nb - 1;
    const int jq = qdiv(j, Ns, rns);
    const int k = j - jq * Ns;
    base[u] = jq * Ns * RADIX + k;
#pragma unroll
    for (int s = 0; s < RADIX; ++s) v[u][s] = buf[j + s * nb];
#pragma unroll
    for (int s = 1; s < RADIX; ++s) w[u][s] = tw_get(tw, s * k * st, inv);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
#pragma unroll
    for (int s = 1; s < RADIX; ++s) v[u][s] = cmul(v[u][s], w[u][s]);
    if (RADIX == 2) dft2(v[u]);
    if (RADIX == 3) dft3(v[u], inv);
    if (RADIX == 4) dft4(v[u], inv);
    if (RADIX == 5) dft5(v[u], inv);
  }
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (lane + 64 * u < nb)
#pragma unroll
      for (int s = 0; s < RADIX; ++s) buf[base[u] + s * Ns] = v[u][s];
}

// All passes of one transform in place by one wavefront (shape checked by
// ip_shape_ok on the host).
template <typename R>
__device__ void fft_run_ip(cplx<R>* buf, const FftShape& sh, const cplx<R>* __restrict__ tw,
                           bool inv, int lane) {
  int Ns = 1;
  for (int p = 0; p < sh.npass; ++p) {
    const int r = sh.radix[p];
    if (r == 4)
      fft_pass_ip<R, 4>(buf, sh.L, Ns, tw, inv, lane);
    else if (r == 2)
      fft_pass_ip<R, 2>(buf, sh.L, Ns, tw, inv, lane);
    else if (r == 3)
      fft_pass_ip<R, 3>(buf, sh.L, Ns, tw, inv, lane);
    else
      fft_pass_ip<R, 5>(buf, sh.L, Ns, tw, inv, lane);
    wave_lds_sync();
    Ns *= r;
  }
}

// ---------------------------------------------------------------------------
// The in-place transform with the length a compile-time constant: the pass
// schedule is make_shape's (radix 4 first, then 2, 3, 5), and every stride,
// chunk count and quotient of a pass folds into the instruction stream.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int ip_radix(int L, int p) {
  const int pref[4] = {4, 2, 3, 5};
  int n = L, q = 0;
  for (int i = 0; i < 4; ++i)
    while (n % pref[i] == 0) {
      if (q == p) return pref[i];
      n /= pref[i];
      ++q;
    }
  return 0;   // no pass p (or a factor above 5)
}

__host__ __device__ constexpr int ip_npass(int L) {
  int q = 0;
  while (ip_radix(L, q) != 0) ++q;
  return q;
}

// Per-pass twiddle layout: pass P (radix R, NS points already combined)
// holds W_L^(s k L/(NS R)) for s = 1..R-1, k = 0..NS-1 at
// ip_twiddle_offset(L, P) + (s-1) NS + k; the passes together use L - 1
// entries.
__host__ __device__ constexpr int ip_twiddle_offset(int L, int P) {
  int off = 0, ns = 1;
  for (int q = 0; q < P; ++q) {
    const int r = ip_radix(L, q);
    off += (r - 1) * ns;
    ns *= r;
  }
  return off;
}

template <typename R, int L, int RADIX, int NS, int TWO>
__device__ __forceinline__ void fft_pass_ct(cplx<R>* buf, const cplx<R>* __restrict__ tw,
                                            bool inv, int lane) {
  constexpr int NB = L / RADIX;                 // butterflies
  constexpr int U = (NB + 63) / 64;             // chunks of 64 butterflies
  static_assert(U <= IpChunk<RADIX>::U, "in-place chunk exceeds the register budget");
  cplx<R> v[U][RADIX], w[U][RADIX];
  int base[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int jr = lane + 64 * u;
    const int j = (NB % 64 == 0 || jr < NB) ? jr : NB - 1;
    const int jq = j / NS;
    const int k = j - jq * NS;
    base[u] = jq * NS * RADIX + k;
#pragma unroll
    for (int s = 0; s < RADIX; ++s) v[u][s] = buf[j + s * NB];
    // the first pass (NS = 1) has only the twiddle W^0 = 1: none is read
    if constexpr (NS > 1)
#pragma unroll
      for (int s = 1; s < RADIX; ++s) w[u][s] = tw_get(tw, TWO + (s - 1) * NS + k, inv);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if constexpr (NS > 1)
#pragma unroll
      for (int s = 1; s < RADIX; ++s) v[u][s] = cmul(v[u][s], w[u][s]);
    if (RADIX == 2) dft2(v[u]);
    if (RADIX == 3) dft3(v[u], inv);
    if (RADIX == 4) dft4(v[u], inv);
    if (RADIX == 5) dft5(v[u], inv);
  }
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (NB % 64 == 0 || lane + 64 * u < NB)
#pragma unroll
      for (int s = 0; s < RADIX; ++s) buf[base[u] + s * NS] = v[u][s];
}

// All passes; tw is the per-pass twiddle table (ip_twiddle_offset).
template <typename R, int L, int P = 0, int NS = 1>
__device__ __forceinline__ void fft_run_ct(cplx<R>* buf, const cplx<R>* __restrict__ tw, bool inv,
                                           int lane) {
  if constexpr (P < ip_npass(L)) {
    constexpr int RADIX = ip_radix(L, P);
    fft_pass_ct<R, L, RADIX, NS, ip_twiddle_offset(L, P)>(buf, tw, inv, lane);
    wave_lds_sync();
    fft_run_ct<R, L, P + 1, NS * RADIX>(buf, tw, inv, lane);
  }
}

}  // namespace sddc
