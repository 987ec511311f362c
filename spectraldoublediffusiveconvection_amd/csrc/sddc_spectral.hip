// Spectral-space kernels: same-parity suffix sums, the fused "spectral stage"
// (Chebyshev contractions on MFMA + theta recurrences + linear terms +
// transposed slab output), right-hand-side assembly, symmetry mask and the
// diagnostics reduction.
#include <type_traits>

#include "sddc_internal.h"

namespace sddc {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// P[r] = sum_{m>r, m = r mod 2} psi[m]: the running sum f_e / b of A2_SINE,
// A2_SINE_R2 and J_theta_RT (Matrix_Operators.py:222-231, 441-450, 502-511).
// One workgroup per instance; threads (parity, i, segment): each segment is
// summed, segment totals are scanned in LDS, then each segment is rewritten.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_suffix(const double* __restrict__ X,
                                                 double* __restrict__ P, int K, int nr,
                                                 int nseg) {
  extern __shared__ double tot[];  // [nseg][2*nr]
  const int b = blockIdx.x;
  const size_t N = (size_t)K * nr;
  const double* psi = X + (size_t)b * 3 * N;
  double* Pb = P + (size_t)b * N;
  const int lanes = 2 * nr;
  const int t = threadIdx.x;
  const int s = t / lanes;           // segment, 0 = highest rows
  const int pi = t - s * lanes;
  const int p = pi / nr, i = pi - p * nr;
  const int top = ((K - 1) % 2 == p) ? K - 1 : K - 2;   // highest row of parity p
  const int cnt = top / 2 + 1;                           // rows of parity p
  const int per = (cnt + nseg - 1) / nseg;
  const int r_hi = top - 2 * per * s;                    // first row of this segment
  int nrows = 0;
  if (s < nseg) {
    nrows = cnt - per * s;
    nrows = nrows < 0 ? 0 : (nrows > per ? per : nrows);
    double acc = 0.0;
    for (int q = 0; q < nrows; ++q) acc += psi[(size_t)(r_hi - 2 * q) * nr + i];
    tot[s * lanes + pi] = acc;
  }
  __syncthreads();
  if (s < nseg) {
    double acc = 0.0;
    for (int u = 0; u < s; ++u) acc += tot[u * lanes + pi];
    for (int q = 0; q < nrows; ++q) {
      const size_t o = (size_t)(r_hi - 2 * q) * nr + i;
      Pb[o] = acc;
      acc += psi[o];
    }
  }
}

int launch_suffix(const Plan& p, int B, const double* X, double* P, hipStream_t s) {
  if (B <= 0) return SDDC_OK;
  const int lanes = 2 * p.nr;
  int nseg = 1024 / lanes;
  const int half = (p.K + 1) / 2;
  const int want = (half + 7) / 8;   // at least 8 rows per segment
  if (nseg > want) nseg = want;
  if (nseg < 1) nseg = 1;
  int threads = ((nseg * lanes + 63) / 64) * 64;
  hipLaunchKernelGGL(k_suffix, dim3(B), dim3(threads), nseg * lanes * sizeof(double), s, X, P,
                     p.K, p.nr, nseg);
  return hip_check(hipGetLastError(), "k_suffix");
}

// ---------------------------------------------------------------------------
// Spectral stage: one workgroup per (tile of 16 modes, run of nb instances).
// The workgroup loads its Chebyshev operator tiles once and walks its
// instances; the state tiles of instance u+1 are loaded into registers while
// instance u is computed.
// ---------------------------------------------------------------------------
constexpr int kMaxSpectralInstances = 4;   // instances per spectral workgroup, at most

struct SpectralArgs {
  const double* X;
  const double* P;
  const double* Ra;
  const double* Ras;
  const double* A64;   // MFMA-tiled Dr | D2r2 | D2s (float64)
  const float* A32;    // MFMA-tiled Dr | D2r2 (float32)
  const double* vecs;
  double* L;           // [b][3][K][nr] or null
  void* slabs;         // [b][7][nr][K] (real_t) or null
  double* theta;       // [b][4][K][nr] or null
  double* slab9;       // [b][9][nr][K] or null
  int K, nr, B, nb;    // nb: instances per workgroup
  double dt, Pr;
};

// Contraction units of the spectral stage's MFMA phase, (contraction << 3) |
// row tile, for wavefront w (of five) and its j-th unit; -1: none.  float64
// plans: wavefront c runs contraction c.  Mixed plans: the float64 units
// (contraction 2) go one per wavefront, first in its list; the float units
// fill the wavefronts up to about equal matrix-core time (a float64 unit
// counts twice).
template <bool kMixed, int NT>
__host__ __device__ constexpr int spectral_units() {
  return !kMixed ? NT : (NT == 1 ? 1 : (NT == 2 ? 3 : (NT == 3 ? 3 : 4)));
}

template <bool kMixed, int NT>
__host__ __device__ constexpr int spectral_unit(int w, int j) {
  if (!kMixed) return j < NT ? (w << 3) | j : -1;
  constexpr int U = spectral_units<kMixed, NT>();
  // rows: wavefronts; entries: (c << 3) | it, -1 = none
  constexpr int t1[5][1] = {{0}, {8}, {16}, {24}, {32}};
  constexpr int t2[5][3] = {{16, -1, -1}, {17, -1, -1}, {0, 1, 8}, {9, 24, 25}, {32, 33, -1}};
  constexpr int t3[5][3] = {{16, 0, 1}, {17, 2, 8}, {18, 9, 10}, {24, 25, 26}, {32, 33, 34}};
  constexpr int t4[5][4] = {{16, 0, 1, 2}, {17, 3, 8, 9}, {18, 10, 11, 24}, {19, 25, 26, 27},
                            {32, 33, 34, 35}};
  if (j >= U) return -1;
  if constexpr (NT == 1) return t1[w][j];
  else if constexpr (NT == 2) return t2[w][j];
  else if constexpr (NT == 3) return t3[w][j];
  else return t4[w][j];
}

// In mixed plans only a wavefront's first unit may be the float64
// contraction; the others are float and load their tiles as floats.
template <int NT>
constexpr bool mixed_units_after_first_are_float() {
  for (int w = 0; w < 5; ++w)
    for (int j = 1; j < spectral_units<true, NT>(); ++j)
      if (spectral_unit<true, NT>(w, j) >= 0 && (spectral_unit<true, NT>(w, j) >> 3) == 2)
        return false;
  return true;
}

// Waves per SIMD the register budget must allow: the mixed 48-row kernel's
// LDS (about 51 KB) fits three workgroups of five wavefronts per CU, which
// needs four waves per SIMD (at most 128 VGPRs); left free, the compiler
// took 129.
template <typename R, int NRP>
constexpr int spectral_waves_per_simd() {
  return std::is_same<R, float>::value && NRP == 48 ? 4 : 1;
}

template <typename R, int NRP>
__global__ __launch_bounds__(320)
__attribute__((amdgpu_waves_per_eu(spectral_waves_per_simd<R, NRP>()))) void k_spectral(
    SpectralArgs a) {
  constexpr int NT = NRP / 16;
  constexpr int NK = NRP / 4;
  // rows padded to an odd number of doubles: the MFMA operand reads take one
  // column of 16 rows, which then fall in different LDS banks
  constexpr int LD = NRP + 1;
  __shared__ double sPsi[kTile + 2][LD];   // modes k0-1 .. k0+16
  __shared__ double sP[kTile + 1][LD];     // modes k0 .. k0+16
  __shared__ double sT[kTile + 1][LD];
  __shared__ double sS[kTile + 1][LD];
  // contraction results [mode][i]: Dr psi, Dr T, Dr S in the precision they
  // are computed in (float in mixed precision: 3 workgroups fit per CU);
  // D2r2 psi and D2s psi in float64 (overwritten by omega and J_theta psi)
  __shared__ R sCr[3][kTile][NRP + 1];
  __shared__ double sCd[2][kTile][NRP + 1];
  constexpr int kVecs = V_GR + 1;           // IR2, IR4, DT0, RSQ, GR
  __shared__ double sV[kVecs * NRP];        // radial vectors
  __shared__ double sRa[2][kMaxSpectralInstances];   // Ra_T, Ra_S of the instances
  const int K = a.K, nr = a.nr;
  const int bfirst = blockIdx.y * a.nb;
  const int nb = min(a.nb, a.B - bfirst);
  const int k0 = blockIdx.x * kTile;
  const size_t N = (size_t)K * nr;

  // Contractions (nr x nr operator times a 16-mode tile), each split into
  // NT row tiles ("units"):
  //  0 Dr psi, 1 D2r2 psi, 2 D2s psi (float64 always: linear term), 3 Dr T, 4 Dr S.
  // Wavefront w runs the units spectral_unit<mixed, NT>(w, j), j < kUnits.
  // In mixed precision the float64 units, which take twice as long on the
  // matrix cores, are spread one per wavefront; unit 0 of a wavefront is the
  // only one that may be float64, so the other units' tiles are held as
  // floats.
  constexpr bool kMixed = std::is_same<R, float>::value;
  constexpr int kUnits = spectral_units<kMixed, NT>();
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int col = lane & 15, kq = lane >> 4;
  double a0[NK];                                    // unit 0 (float64 or float)
  float af[kUnits > 1 ? kUnits - 1 : 1][NK];        // units 1.. (mixed: float)
  double ad[kMixed ? 1 : (kUnits > 1 ? kUnits - 1 : 1)][NK];   // units 1.. (float64 plan)
  static_assert(!kMixed || mixed_units_after_first_are_float<NT>(),
                "mixed plans hold units after the first as floats");
  // Whether a unit is float64 depends on the wavefront index, known only at
  // run time, so the loads carry no per-unit choice of array or conversion
  // (either would make each load wait before the next is issued): unit 0
  // always loads its float64 tile (rounded to float where it is used, which
  // gives the float array's value: that array holds the same operators
  // rounded to float), the float units load the float array.
#pragma unroll
  for (int j = 0; j < kUnits; ++j) {
    const int code = spectral_unit<kMixed, NT>(w, j);
    if (code < 0) continue;
    const int cc = code >> 3, it = code & 7;
    const int slot = (cc == 1) ? 1 : (cc == 2 ? 2 : 0);   // Dr, D2r2, D2s
    const size_t base = (size_t)slot * NT * NK * 64 + (size_t)it * NK * 64 + lane;
#pragma unroll
    for (int kk = 0; kk < NK; ++kk) {
      if (j == 0)
        a0[kk] = a.A64[base + kk * 64];
      else if (kMixed)
        af[j - 1][kk] = a.A32[base + kk * 64];
      else
        ad[kMixed ? 0 : j - 1][kk] = a.A64[base + kk * 64];
    }
  }
  constexpr int kVecIters = (kVecs * NRP + 319) / 320;
#pragma unroll
  for (int it = 0; it < kVecIters; ++it) {
    const int e = threadIdx.x + 320 * it;
    if (e < kVecs * NRP) sV[e] = a.vecs[e];
  }
  // the parameters are staged with the prologue: a global load inside the
  // loop would be waited for together with the next instance's prefetch
  if (threadIdx.x < 2 * kMaxSpectralInstances) {
    const int u = threadIdx.x % kMaxSpectralInstances, q = threadIdx.x / kMaxSpectralInstances;
    const double* src = q == 0 ? a.Ra : a.Ras;
    sRa[q][u] = (src && u < nb) ? src[bfirst + u] : 0.0;
  }

  // state tiles of one instance, held in registers between their load and
  // their store to LDS: psi (modes k0-1 ..), P, T, S (modes k0 ..)
  constexpr int kLoadIters = ((kTile + 2) * NRP + 319) / 320;
  constexpr int kOutIters = (kTile * NRP + 319) / 320;   // outputs per thread, at most
  double st[kLoadIters][4];
  auto fetch = [&](int b, int tx) {
    const double* Xb = a.X + (size_t)b * 3 * N;
    const double* Pb = a.P + (size_t)b * N;
#pragma unroll
    for (int it = 0; it < kLoadIters; ++it) {
      const int idx = tx + 320 * it;
      const int t = idx / NRP, i = idx - t * NRP;
      const int k = k0 - 1 + t;
      const bool ok = idx < (kTile + 2) * NRP && k >= 0 && k < K && i < nr;
      // every lane loads (entry 0 where out of range; the store to LDS
      // writes zero there): no branch or select waits for these loads,
      // which stay in flight until their values are stored
      const size_t o = ok ? (size_t)k * nr + i : 0;
      st[it][0] = Xb[o];
      st[it][1] = Pb[o];
      st[it][2] = Xb[N + o];
      st[it][3] = Xb[2 * N + o];
    }
  };
  auto put = [&](int cc, int rl, int i, double v) {
    if (cc == 1)
      sCd[0][rl][i] = v;
    else if (cc == 2)
      sCd[1][rl][i] = v;
    else
      sCr[cc == 0 ? 0 : cc - 2][rl][i] = (R)v;   // exact: a value of type R
  };
  const double* vIR2 = sV + V_IR2 * NRP;
  const double* vIR4 = sV + V_IR4 * NRP;
  const double* vDT0 = sV + V_DT0 * NRP;
  const double* vRSQ = sV + V_RSQ * NRP;
  const double* vGR = sV + V_GR * NRP;
  const double dtpr = a.dt * a.Pr;

  fetch(bfirst, threadIdx.x);
  // the operator tiles are complete before the loop: the compiler's wait
  // counts inside it then track only the loads of the next instance
  drain_loads();
  for (int u = 0; u < nb; ++u) {
    const int b = bfirst + u;
    // the thread index, opaque to the optimiser: index arithmetic is redone
    // per instance instead of being held in registers across the loop
    int tx = threadIdx.x;
    asm volatile("" : "+v"(tx));
    if (u > 0) lds_barrier();   // the previous instance is done with every LDS array
#pragma unroll
    for (int it = 0; it < kLoadIters; ++it) {
      const int idx = tx + 320 * it;
      if (idx < (kTile + 2) * NRP) {
        const int t = idx / NRP, i = idx - t * NRP;
        const int k = k0 - 1 + t;
        const bool ok = k >= 0 && k < K && i < nr;
        sPsi[t][i] = ok ? st[it][0] : 0.0;
        if (t >= 1) {
          sP[t - 1][i] = ok ? st[it][1] : 0.0;
          sT[t - 1][i] = ok ? st[it][2] : 0.0;
          sS[t - 1][i] = ok ? st[it][3] : 0.0;
        }
      }
    }
    if (u + 1 < nb) fetch(b + 1, tx);
    lds_barrier();

    // (nr x nr) @ (nr x 16 modes) on MFMA.  Both shapes take A[i][k] from lane
    // i + 16k and B[k][j] from lane j + 16k; their output layouts differ.
    {
      // the tile registers are opaque to the optimiser here, so that it does
      // not keep converted copies of them live across the instance loop
#pragma unroll
      for (int kk = 0; kk < NK; ++kk) asm volatile("" : "+v"(a0[kk]));
#pragma unroll
      for (int j = 0; j < kUnits; ++j) {
        const int code = spectral_unit<kMixed, NT>(w, j);
        if (code < 0) continue;
        const int cc = code >> 3, it = code & 7;
        const double* src = (cc <= 2) ? &sPsi[1][0] : (cc == 3 ? &sT[0][0] : &sS[0][0]);
        const double* srow = src + col * LD + kq;
        if (!kMixed || cc == 2) {
          v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int kk = 0; kk < NK; ++kk) {
            const double av = (j == 0) ? a0[kk] : ad[kMixed ? 0 : j - 1][kk];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, srow[4 * kk], acc, 0, 0, 0);
          }
          // float64 16x16x4 output: lane l, register v holds row l/16 + 4v
#pragma unroll
          for (int v = 0; v < 4; ++v) put(cc, col, 16 * it + kq + 4 * v, acc[v]);
        } else {
          v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int kk = 0; kk < NK; ++kk) {
            const float av = (j == 0) ? (float)a0[kk] : af[j > 0 ? j - 1 : 0][kk];
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, (float)srow[4 * kk], acc, 0, 0, 0);
          }
          // float32 16x16x4 output: lane l, register v holds row 4*(l/16) + v
#pragma unroll
          for (int v = 0; v < 4; ++v) put(cc, col, 16 * it + 4 * kq + v, (double)acc[v]);
        }
      }
    }
    lds_barrier();

    const double Ra = sRa[0][u];
    const double Ras = sRa[1][u];
    double* Lb = a.L ? a.L + (size_t)b * 3 * N : nullptr;
    double* Th = a.theta ? a.theta + (size_t)b * 4 * N : nullptr;
#pragma unroll
    for (int it = 0; it < kOutIters; ++it) {
      const int idx = tx + 320 * it;
      const int rl = idx / nr, i = idx - rl * nr;
      const int r = k0 + rl;
      if (idx >= kTile * nr || r >= K) continue;
      const double psi = sPsi[rl + 1][i];
      const double j = r + 1.0;                        // sine wavenumber of row r
      const double t = (j + 1.0) * psi + 2.0 * sP[rl][i];
      // A2_SINE_R2 (Matrix_Operators.py:510) and A2_SINE (:230)
      const double Om = sCd[0][rl][i] - j * (vIR4[i] * t);
      const double A2 = sCd[1][rl][i] - (j * vIR2[i]) * t;
      // J_theta_RT (Matrix_Operators.py:449-455): b = inclusive same-parity sum from row r+1
      const double Q = sPsi[rl + 2][i] + sP[rl + 1][i];
      const double JP = (r == 0) ? 1.0 * Q : (r + 1.0) * sPsi[rl][i] + 2.0 * Q;
      const size_t o = (size_t)r * nr + i;
      const double psiT0 = vDT0[i] * JP;               // DT0_theta (Matrix_Operators.py:170-172)
      if (Lb) {
        // Step_Python right-hand sides (Main.py:271, 276, 280)
        double grv = 0.0;
        if (r <= K - 2) {
          const double v = Ra * sT[rl + 1][i] - Ras * sS[rl + 1][i];
          grv = ((double)(-(r + 1)) * vGR[i]) * v;     // kGR_RT row r <- cosine row r+1
        }
        Lb[o] = A2 + dtpr * grv;
        Lb[N + o] = vRSQ[i] * sT[rl][i] - a.dt * psiT0;
        Lb[2 * N + o] = vRSQ[i] * sS[rl][i] - a.dt * psiT0;
      }
      if (Th) {
        Th[o] = JP;
        Th[N + o] = psiT0;
        Th[2 * N + o] = A2;
        Th[3 * N + o] = Om;
      }
      sCd[0][rl][i] = Om;   // same thread read these entries above
      sCd[1][rl][i] = JP;
    }
    lds_barrier();

    if (a.slabs) {
      const size_t nK = (size_t)nr * K;
      R* S7 = reinterpret_cast<R*>(a.slabs) + (size_t)b * 7 * nK;
#pragma unroll
      for (int it = 0; it < kOutIters; ++it) {
        const int idx = tx + 320 * it;
        const int i = idx / kTile, rl = idx - i * kTile;
        const int r = k0 + rl;
        if (idx >= nr * kTile || r >= K) continue;
        const size_t row = (size_t)i * K;
        S7[row + r] = (R)sCd[1][rl][i];
        S7[3 * nK + row + r] = (R)sCr[1][rl][i];
        S7[4 * nK + row + r] = (R)sCr[2][rl][i];
        S7[5 * nK + row + r] = (R)sT[rl][i];
        S7[6 * nK + row + r] = (R)sS[rl][i];
        if (r + 1 < K) {   // sine rows move to the sinusoid index r+1 (Matrix_Operators.py:734-737)
          S7[nK + row + r + 1] = (R)sCr[0][rl][i];
          S7[2 * nK + row + r + 1] = (R)sCd[0][rl][i];
        }
        if (r == 0) {
          S7[nK + row] = (R)0;
          S7[2 * nK + row] = (R)0;
        }
      }
    }
    if (a.slab9) {
      const size_t nK = (size_t)nr * K;
      double* S9 = a.slab9 + (size_t)b * 9 * nK;
#pragma unroll
      for (int it = 0; it < kOutIters; ++it) {
        const int idx = tx + 320 * it;
        const int i = idx / kTile, rl = idx - i * kTile;
        const int r = k0 + rl;
        if (idx >= nr * kTile || r >= K) continue;
        const size_t row = (size_t)i * K;
        // order: JT, kDpsi, komega, DT, DC, omega, Dpsi, kT, kC (Matrix_Operators.py:740)
        S9[row + r] = sCd[1][rl][i];
        S9[3 * nK + row + r] = sCr[1][rl][i];
        S9[4 * nK + row + r] = sCr[2][rl][i];
        S9[7 * nK + row + r] = (double)(-r) * sT[rl][i];
        S9[8 * nK + row + r] = (double)(-r) * sS[rl][i];
        if (r + 1 < K) {
          const double ks = r + 1.0;
          S9[1 * nK + row + r + 1] = ks * sCr[0][rl][i];
          S9[2 * nK + row + r + 1] = ks * sCd[0][rl][i];
          S9[5 * nK + row + r + 1] = sCd[0][rl][i];
          S9[6 * nK + row + r + 1] = sCr[0][rl][i];
        }
        if (r == 0) {
          S9[1 * nK + row] = 0.0;
          S9[2 * nK + row] = 0.0;
          S9[5 * nK + row] = 0.0;
          S9[6 * nK + row] = 0.0;
        }
      }
    }
  }
}

// Instances per spectral workgroup.  A workgroup overlaps the loads of its
// next instance with the work on the current one, so runs of up to four
// instances pay off once there are enough (tile, instance) pairs to keep
// every CU busy; measured on the 512-mode, 39-point class at 64 and 256
// instances: four per workgroup beat one, two and eight.
static int spectral_instances(int tiles, int B) {
  const long nb = ((long)tiles * B) / 512;
  return nb < 1 ? 1 : (nb > kMaxSpectralInstances ? kMaxSpectralInstances : (int)nb);
}

template <typename R>
static int launch_spectral_r(const Plan& p, SpectralArgs a, int B, hipStream_t s) {
  const int tiles = (p.K + kTile - 1) / kTile;
  a.B = B;
  a.nb = p.tune_spectral_nb > 0 ? p.tune_spectral_nb : spectral_instances(tiles, B);
  const dim3 grid(tiles, (B + a.nb - 1) / a.nb);
  switch (p.NRP) {
    case 16: hipLaunchKernelGGL((k_spectral<R, 16>), grid, dim3(320), 0, s, a); break;
    case 32: hipLaunchKernelGGL((k_spectral<R, 32>), grid, dim3(320), 0, s, a); break;
    case 48: hipLaunchKernelGGL((k_spectral<R, 48>), grid, dim3(320), 0, s, a); break;
    case 64: hipLaunchKernelGGL((k_spectral<R, 64>), grid, dim3(320), 0, s, a); break;
    default: return fail(SDDC_ERR_UNSUPPORTED, "unsupported radial size");
  }
  return hip_check(hipGetLastError(), "k_spectral");
}

int launch_spectral(const Plan& p, int B, const double* X, const double* P, const double* Ra,
                    const double* Ras, int want_linear, double* L, void* slabs,
                    int slab_precision, double* theta_out, double* slab9, hipStream_t s) {
  if (B <= 0) return SDDC_OK;
  SpectralArgs a;
  a.X = X;
  a.P = P;
  a.Ra = Ra;
  a.Ras = Ras;
  a.A64 = p.mfma64;
  a.A32 = p.mfma32;
  a.vecs = p.vecs;
  a.L = want_linear ? L : nullptr;
  a.slabs = slabs;
  a.theta = theta_out;
  a.slab9 = slab9;
  a.K = p.K;
  a.nr = p.nr;
  a.dt = p.dt;
  a.Pr = p.Pr;
  return slab_precision == SDDC_MIXED ? launch_spectral_r<float>(p, a, B, s)
                                      : launch_spectral_r<double>(p, a, B, s);
}

// ---------------------------------------------------------------------------
// Right-hand side: G = L - dt*F with F_psi = Dr @ F_Q - F_R shifted back to
// code rows (Matrix_Operators.py:791,797,802; Main.py:262,271,276,280).
// ---------------------------------------------------------------------------
struct RhsArgs {
  const void* F;
  const double* L;
  const double* A64;
  const float* A32;
  double* G;
  int K, nr, have_nonlin;
  double scale;   // -dt for the step (NX = -1*dt*FX, Main.py:262), 1 for F itself
};

// Mode tile of a workgroup, grouped by XCD.  Workgroups are dispatched to
// the eight XCDs round-robin, so with tiles in blockIdx order neighbouring
// tiles land on different XCDs; each XCD's L2 then fetches whole 128-byte
// lines of the [row][mode] float slabs for the 64 bytes its tile uses.  With
// the tile count a multiple of 8, XCD c gets the consecutive tiles
// c*T/8 .. (c+1)*T/8 - 1 instead, and the lines its tiles share are fetched
// once (the rhs stage's product reads: 2.5x -> about 1x their bytes).
__device__ __forceinline__ int xcd_grouped_tile(int x, int tiles) {
  return (tiles % 8 == 0) ? (x % 8) * (tiles / 8) + x / 8 : x;
}

// Waves per SIMD the mixed 48-row kernel's registers must allow: its tile
// loads all in flight took it to 76 VGPRs (six workgroups per CU); bounded to
// 72 (70) it keeps seven, 16.8 -> 16.4 us per 64-instance launch, same bits.
// The 64-row kernel would spill under the same bound.
template <typename R, int NRP>
constexpr int rhs_waves_per_simd() {
  return std::is_same<R, float>::value && NRP == 48 ? 7 : 1;
}

template <typename R, int NRP>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(rhs_waves_per_simd<R, NRP>()))) void k_rhs(RhsArgs a) {
  constexpr int NT = NRP / 16;   // <= 4: one row tile per wavefront
  constexpr int NK = NRP / 4;
  // rows padded to an odd length: the transposing fill below and the MFMA
  // operand reads take one column of 16 rows, which then fall in different
  // LDS banks
  __shared__ R sQ[kTile][NRP + 1];
  __shared__ R sR[kTile][NRP + 1];
  __shared__ R sTT[kTile][NRP + 1];
  __shared__ R sSS[kTile][NRP + 1];
  __shared__ R sDQ[kTile][NRP + 1];
  const int K = a.K, nr = a.nr, b = blockIdx.y;
  const int k0 = xcd_grouped_tile(blockIdx.x, gridDim.x) * kTile;
  const size_t N = (size_t)K * nr;
  const size_t nK = (size_t)nr * K;
  const double* Lb = a.L ? a.L + (size_t)b * 3 * N : nullptr;
  double* Gb = a.G + (size_t)b * 3 * N;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int col = lane & 15, kq = lane >> 4;
  const bool mfma = a.have_nonlin && wave < NT;
  // Every global load is issued before any is waited for: the Dr row tile of
  // this wavefront, the linear parts L of this thread's outputs, the product
  // tiles.
  double a64[NK];
  float a32[NK];
  if (mfma) {
    if (std::is_same<R, double>::value) {
      const double* At = a.A64 + ((size_t)wave * NK) * 64 + lane;
#pragma unroll
      for (int kk = 0; kk < NK; ++kk) a64[kk] = At[kk * 64];
    } else {
      const float* At = a.A32 + ((size_t)wave * NK) * 64 + lane;
#pragma unroll
      for (int kk = 0; kk < NK; ++kk) a32[kk] = At[kk * 64];
    }
  }
  constexpr int kOut = (kTile * NRP + 255) / 256;
  double lv[kOut][3];
#pragma unroll
  for (int j = 0; j < kOut; ++j) {
    const int idx = threadIdx.x + 256 * j;
    const int rl = idx / nr, i = idx - rl * nr;
    const bool ok = Lb && idx < kTile * nr && k0 + rl < K;
    const size_t o = ok ? (size_t)(k0 + rl) * nr + i : 0;
#pragma unroll
    for (int f = 0; f < 3; ++f) lv[j][f] = ok ? Lb[f * N + o] : 0.0;
  }
  if (a.have_nonlin) {
    const R* Fb = reinterpret_cast<const R*>(a.F) + (size_t)b * 4 * nK;
    constexpr int kIters = (NRP * kTile + 255) / 256;
    // every lane loads (entry 0 where out of range) before any value is
    // selected or stored: a load under its range condition would be waited
    // for before the next one is issued
    R fv[kIters][4];
#pragma unroll
    for (int it = 0; it < kIters; ++it) {
      const int idx = threadIdx.x + 256 * it;
      const int i = idx / kTile, rl = idx - i * kTile;
      const int r = k0 + rl, m = r + 1;
      const bool in = idx < NRP * kTile && i < nr;
      const size_t row = (size_t)i * K;
      const size_t om = (in && m < K) ? row + m : 0, orr = (in && r < K) ? row + r : 0;
      fv[it][0] = Fb[om];
      fv[it][1] = Fb[nK + om];
      fv[it][2] = Fb[2 * nK + orr];
      fv[it][3] = Fb[3 * nK + orr];
    }
#pragma unroll
    for (int it = 0; it < kIters; ++it) {
      const int idx = threadIdx.x + 256 * it;
      if (idx < NRP * kTile) {
        const int i = idx / kTile, rl = idx - i * kTile;
        const int r = k0 + rl, m = r + 1;
        const bool okm = (i < nr && m < K), okr = (i < nr && r < K);
        sQ[rl][i] = okm ? fv[it][0] : (R)0;
        sR[rl][i] = okm ? fv[it][1] : (R)0;
        sTT[rl][i] = okr ? fv[it][2] : (R)0;
        sSS[rl][i] = okr ? fv[it][3] : (R)0;
      }
    }
    __syncthreads();
    if (mfma) {
      const R* srow = &sQ[col][kq];
      if (std::is_same<R, double>::value) {
        v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < NK; ++kk)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a64[kk], (double)srow[4 * kk], acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) sDQ[col][16 * wave + kq + 4 * v] = (R)acc[v];
      } else {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < NK; ++kk)
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a32[kk], (float)srow[4 * kk], acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) sDQ[col][16 * wave + 4 * kq + v] = (R)acc[v];
      }
    }
    __syncthreads();
  }
  const double sc = a.scale;
#pragma unroll
  for (int j = 0; j < kOut; ++j) {
    const int idx = threadIdx.x + 256 * j;
    const int rl = idx / nr, i = idx - rl * nr;
    const int r = k0 + rl;
    if (idx >= kTile * nr || r >= K) continue;
    const size_t o = (size_t)r * nr + i;
    if (a.have_nonlin) {
      // code row r of the psi term is sinusoid index r+1; row K-1 is zero
      const double Fpsi = (r <= K - 2) ? (double)(sDQ[rl][i] - sR[rl][i]) : 0.0;
      const double Ft = (double)sTT[rl][i], Fs = (double)sSS[rl][i];
      if (Lb) {
        Gb[o] = sc * Fpsi + lv[j][0];
        Gb[N + o] = sc * Ft + lv[j][1];
        Gb[2 * N + o] = sc * Fs + lv[j][2];
      } else {
        Gb[o] = sc * Fpsi;
        Gb[N + o] = sc * Ft;
        Gb[2 * N + o] = sc * Fs;
      }
    } else {
      Gb[o] = lv[j][0];
      Gb[N + o] = lv[j][1];
      Gb[2 * N + o] = lv[j][2];
    }
  }
}

template <typename R>
int launch_rhs_r(const Plan& p, const RhsArgs& a, int B, hipStream_t s) {
  const dim3 grid((p.K + kTile - 1) / kTile, B);
  switch (p.NRP) {
    case 16: hipLaunchKernelGGL((k_rhs<R, 16>), grid, dim3(256), 0, s, a); break;
    case 32: hipLaunchKernelGGL((k_rhs<R, 32>), grid, dim3(256), 0, s, a); break;
    case 48: hipLaunchKernelGGL((k_rhs<R, 48>), grid, dim3(256), 0, s, a); break;
    case 64: hipLaunchKernelGGL((k_rhs<R, 64>), grid, dim3(256), 0, s, a); break;
    default: return fail(SDDC_ERR_UNSUPPORTED, "unsupported radial size");
  }
  return hip_check(hipGetLastError(), "k_rhs");
}

static int rhs_common(const Plan& p, int B, const void* F, const double* L, int have_nonlin,
                      double scale, double* G, hipStream_t s) {
  if (B <= 0) return SDDC_OK;
  RhsArgs a;
  a.F = F;
  a.L = L;
  a.A64 = p.mfma64;
  a.A32 = p.mfma32;
  a.G = G;
  a.K = p.K;
  a.nr = p.nr;
  a.have_nonlin = have_nonlin;
  a.scale = scale;
  return p.precision == SDDC_MIXED ? launch_rhs_r<float>(p, a, B, s)
                                   : launch_rhs_r<double>(p, a, B, s);
}

int launch_rhs(const Plan& p, int B, const void* F, const double* L, int have_nonlin, double* G,
               hipStream_t s) {
  return rhs_common(p, B, F, L, have_nonlin, -1 * p.dt, G, s);
}

int launch_rhs_scaled(const Plan& p, int B, const void* F, const double* L, double scale,
                      double* G, hipStream_t s) {
  return rhs_common(p, B, F, L, 1, scale, G, s);
}

// ---------------------------------------------------------------------------
// Equatorial symmetry mask (Eq_SYM, Main.py:137-176): psi rows with even k and
// T/S rows with odd k are zeroed.
// ---------------------------------------------------------------------------
__global__ void k_mask(const double* __restrict__ X, double* __restrict__ out, int K, int nr,
                       size_t total) {
  const size_t N = (size_t)K * nr;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t w = e % (3 * N);
    const int f = (int)(w / N);
    const int k = (int)((w - f * N) / nr);
    const bool zero = (f == 0) ? (k % 2 == 0) : (k % 2 == 1);
    out[e] = zero ? 0.0 : X[e];
  }
}

int launch_mask(const Plan& p, int B, const double* X, double* out, hipStream_t s) {
  const size_t total = (size_t)B * 3 * p.K * p.nr;
  if (total == 0) return SDDC_OK;
  const int blocks = (int)((total + 255) / 256 < 4096 ? (total + 255) / 256 : 4096);
  hipLaunchKernelGGL(k_mask, dim3(blocks), dim3(256), 0, s, X, out, p.K, p.nr, total);
  return hip_check(hipGetLastError(), "k_mask");
}

// ---------------------------------------------------------------------------
// Diagnostics: ||X||_2, kinetic energy (radial trapezoid over the interior
// points of the per-row theta integrals, Main.py:129-134) and Nu-1 of T and S
// at R_1 and at R_2 (Main.py:52-68; the reference returns the R_1 value and
// checks it against R_2).  One workgroup per instance; out[b] holds
// (norm, KE, NuT(R_1), NuS(R_1), NuT(R_2), NuS(R_2)).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_diag_final(const double* __restrict__ X,
                                                    const double* __restrict__ rowval,
                                                    const double* __restrict__ Dfull,
                                                    const double* __restrict__ Rg, int K, int nr,
                                                    double d, double* __restrict__ out) {
  constexpr int NV = 6;
  __shared__ double red[NV][256];
  const int b = blockIdx.x;
  const size_t N = (size_t)K * nr;
  const double* Xb = X + (size_t)b * 3 * N;
  const int Np1 = nr + 2;
  const double* Dlast = Dfull + (size_t)(Np1 - 1) * Np1;   // row of R_2
  double v[NV] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (size_t e = threadIdx.x; e < 3 * N; e += blockDim.x) v[0] += Xb[e] * Xb[e];
  // Nusselt: sum over even k of (D[row,1:-1] . T_k) / (1 - k^2), rows 0 and N_r
  for (int k = 2 * threadIdx.x; k < K; k += 2 * blockDim.x) {
    double aT = 0.0, aS = 0.0, bT = 0.0, bS = 0.0;
    for (int i = 0; i < nr; ++i) {
      const double t = Xb[N + (size_t)k * nr + i], c = Xb[2 * N + (size_t)k * nr + i];
      aT += Dfull[i + 1] * t;
      aS += Dfull[i + 1] * c;
      bT += Dlast[i + 1] * t;
      bS += Dlast[i + 1] * c;
    }
    const double den = 1.0 - (double)k * (double)k;
    v[2] += aT / den;
    v[3] += aS / den;
    v[4] += bT / den;
    v[5] += bS / den;
  }
  for (int i = threadIdx.x; i < nr; i += blockDim.x) {
    const double* x = Rg + 1;   // interior points
    double w;
    if (nr == 1) w = 0.0;
    else if (i == 0) w = 0.5 * (x[1] - x[0]);
    else if (i == nr - 1) w = 0.5 * (x[nr - 1] - x[nr - 2]);
    else w = 0.5 * (x[i + 1] - x[i - 1]);
    if (rowval) v[1] += w * rowval[(size_t)b * nr + i];
  }
  for (int q = 0; q < NV; ++q) red[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int q = 0; q < NV; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double R1 = 1 / d, R2 = (1 + d) / d;
    const double A_T = (R1 * R2) / (R1 - R2);
    const double V = (2.0 / 3.0) * (Rg[Np1 - 1] * Rg[Np1 - 1] * Rg[Np1 - 1] - Rg[0] * Rg[0] * Rg[0]);
    const double c0 = (Rg[0] * Rg[0]) / A_T, c1 = (Rg[Np1 - 1] * Rg[Np1 - 1]) / A_T;
    double* o = out + (size_t)b * NV;
    o[0] = sqrt(red[0][0]);
    o[1] = (0.5 / V) * red[1][0];
    o[2] = c0 * red[2][0];
    o[3] = c0 * red[3][0];
    o[4] = c1 * red[4][0];
    o[5] = c1 * red[5][0];
  }
}

int launch_diag_final(const Plan& p, int B, const double* X, const double* rowval, double* out,
                      hipStream_t s) {
  if (B <= 0) return SDDC_OK;
  hipLaunchKernelGGL(k_diag_final, dim3(B), dim3(256), 0, s, X, rowval, p.Dfull, p.R, p.K, p.nr,
                     p.d, out);
  return hip_check(hipGetLastError(), "k_diag_final");
}

}  // namespace sddc
