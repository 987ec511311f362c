// Implicit solves of the step, float64 always: block back substitution over
// theta modes with pre-inverted per-mode radial blocks.
//   A4_BSub_TSTEP_V2   (Matrix_Operators.py:1115-1194)  field 0 (psi)
//   NAB2_BSub_TSTEP_V2 (Matrix_Operators.py:1033-1086)  fields 1 (T), 2 (S)
//
// Each field has two independent chains (the two row parities).  A chain is a
// linear recurrence over its rows, high to low: a state vector (NAB2: the
// running sum b, nr values; A4: the running sums f_e and bf_e, 2 nr values)
// is updated after every row, and the row's solution is Linv_row applied to
// g_row plus terms in the state.  The chain is cut into segments of m rows.
// The recurrence is affine in the state, so
//   * the solution of a row is the one computed with the segment's entry state
//     set to zero, plus Q_row times the true entry state;
//   * a segment's exit state is its zero-entry exit plus E_seg times its entry
//     state.
// Q and E depend on the operators only and are built once per plan.  A solve
// is two launches:
//   k_chain_local  every segment of every chain at once, zero entry state;
//   k_chain_join   entry states by the short recurrence over segments, then
//                  the Q corrections of the rows, which are independent.
// The sequential depth of a chain drops from K/2 rows to m + K/(2m).
//
// Work layout: a workgroup holds NRP/16 wavefronts; wavefront w owns radial
// rows 16w..16w+15 with four lanes per row (a quad): lane l has row
// 16w + l/4 and the columns c = l (mod 4) of that row.  Matrices are stored
// quad-packed, so that every 16-byte load instruction of a wavefront reads
// 1 KiB of consecutive memory: element (r, c) of an NRP x NC matrix sits at
//   ((w*(NC/8) + q)*64 + l)*2 + h,   r = 16w + l/4,  c = l%4 + 8q + 4h.
// Rows r >= nr are padding: their matrix columns are zero, their lanes read
// the operator entries and inputs of row nr-1 (cache lines the valid lanes
// fetch anyway), so their values stay finite, never reach a real row through
// the zero columns, and are not stored: the NAB2 chains and the
// two-exchange A4 chains of the local launch (local_nab2, local_a4x) mask
// those lanes off their stores; the one-exchange A4 chain (local_a4), the
// joins and the MFMA kernel send them to a sink.  Those two chains also run
// their rows in full groups of prefetched rows with no per-row liveness
// selects (the last, partial group checks), and address every load and
// store as a wave-uniform row base plus a fixed 32-bit lane offset.  That
// form fits the compiler's scalar-base addressing, so a row spends no
// vector instructions on addresses, and the width-5 kernel needs 140
// instead of 178 VGPRs (three wavefronts per SIMD instead of two).  In
// local_a4 only the addressing pays (masked stores and peeled groups made
// its rows slower), and the joins need more registers either way, so those
// keep the sink form.
// A matrix-vector product is one partial dot product per lane (NC/4 terms)
// and a quad reduction (two DPP exchanges, which leave the same sum in all
// four lanes).  The vector is broadcast through LDS, stored as [c%4][c/4] so
// each lane reads its NC/4 entries contiguously.  Operator data of the next
// kDepth rows is loaded into registers ahead of use.
#include <cmath>
#include <cstdlib>
#include <vector>

#include "sddc_internal.h"

namespace sddc {

constexpr int kDepth = 2;    // chain rows of operator data in flight per lane
constexpr int kDepthE = 2;   // segment exit maps in flight per lane (join)
constexpr int kDepthM = 4;   // rows in flight in the MFMA kernel (few waves: deep prefetch)

struct ChainArgs {
  const double* G;          // right-hand sides, field f of instance b at b*gstride+goff[f]
  double* X;                // solutions, b*xstride + xoff[f]
  double* P;                // psi suffix sums [B][N] or null
  const double* Lq[3];      // packed block inverses per field, block index jj = K-1-row
  const double* DLq;        // packed D2 Linv_j of the A4 solve, per mode
  const double* D2q;        // packed D2 of the A4 solve (one block, all modes)
  const double* Qq[3];      // packed corrections, index jj
  const double* Eq[3];      // packed segment exit maps, index chain*nseg + seg
  const double* Lm[3];      // MFMA-packed block inverses (16-instance kernel)
  const double* DLm;        // MFMA-packed D2 Linv_j (16-instance kernel)
  const double* vecs;
  double* EX;               // zero-entry exit states [b][3][2][nseg][2*NRP]
  double* scratch;          // 2N + 256 doubles: sink of stores of padded instances/rows
  long long gstride, xstride;
  long long goff[3], xoff[3];
  int K, nr, B, m, nseg, symmetric, fields_mask;
  double dtf[3];            // Pr*dt, dt, Tau*dt
};

template <int NRS>
struct CShape {
  static constexpr int NRP = ((NRS + 15) / 16) * 16;
  static constexpr int WR = NRP / 16;     // wavefronts per workgroup
  static constexpr int NQ = NRS / 8;      // 16-byte loads per lane of an NRP x NRS block
  static constexpr int CPL = NRS / 4;     // entries of an NRS vector per lane
  static constexpr int BLK = NRP * NRS;   // doubles per packed NRP x NRS block
};

// Position of vector entry c in the LDS broadcast layout of a vector with
// 4*cpl entries.
__device__ __forceinline__ int vpos(int c, int cpl) { return (c & 3) * cpl + (c >> 2); }

template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Sum over the four lanes of a quad; every lane gets the same value.
__device__ __forceinline__ double quad_sum(double v) {
  v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]
  return v;
}

template <int NQ>
__device__ __forceinline__ void load_blk(double2 (&a)[NQ], const double* p) {
#pragma unroll
  for (int q = 0; q < NQ; ++q) a[q] = *reinterpret_cast<const double2*>(p + q * 128);
}

// Partial dot product of the lane's packed matrix entries with its vector
// entries x[0 .. 2NQ) (column l%4 + 8q + 4h pairs with x[2q + h]).
template <int NQ>
__device__ __forceinline__ double pdot(const double2 (&a)[NQ], const double* x) {
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const double2 v = *reinterpret_cast<const double2*>(x + 2 * q);
    s0 = fma(a[q].x, v.x, s0);
    s1 = fma(a[q].y, v.y, s1);
  }
  return s0 + s1;
}

template <int NQ>
__device__ __forceinline__ double pdot_reg(const double2 (&a)[NQ], const double (&x)[2 * NQ]) {
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    s0 = fma(a[q].x, x[2 * q], s0);
    s1 = fma(a[q].y, x[2 * q + 1], s1);
  }
  return s0 + s1;
}

// A wave-uniform pointer plus a lane's byte offset: the compiler can then
// address with a scalar base and one 32-bit vector offset, so a loop that
// moves the base from row to row spends no vector instructions on addresses.
__device__ __forceinline__ const double* lane_at(const double* base, unsigned byte_off) {
  return reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + byte_off);
}
__device__ __forceinline__ double* lane_at(double* base, unsigned byte_off) {
  return reinterpret_cast<double*>(reinterpret_cast<char*>(base) + byte_off);
}

// A loop-invariant lane offset made opaque to loop-invariant code motion:
// otherwise the compiler adds it to each uniform base once, outside the
// loop, as a 64-bit vector address that every row then advances with vector
// adds.  No instruction is emitted.
__device__ __forceinline__ unsigned opaque(unsigned v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <int NB>
struct Inst {
  const double* g[NB];
  double* x[NB];
  double* p[NB];
  double* ex[NB];   // exit states of segment 0 of this (field, chain)
};

// The instances of instance group `group` (NB instances per workgroup).
template <int NB>
__device__ __forceinline__ Inst<NB> instances(const ChainArgs& a, int field, int chain,
                                              int group) {
  Inst<NB> in;
  const size_t N = (size_t)a.K * a.nr;
  const size_t exs = (size_t)a.nseg * 2 * (((a.nr + 15) / 16) * 16);
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    const int b = group * NB + t;
    const bool valid = b < a.B;
    const int bc = valid ? b : a.B - 1;
    in.g[t] = a.G + (size_t)bc * a.gstride + a.goff[field];
    // instances past the batch write into scratch, so every wavefront runs
    // the same instructions.  Their segment exit states go to those of the
    // last instance, whose inputs they read and whose values they repeat
    // bit for bit: the exit-state workspace holds the batch's instances
    // only, and with three instances per workgroup the padded instances
    // would reach past it (or, in a sub-batch, into the next sub-batch's).
    in.x[t] = valid ? a.X + (size_t)b * a.xstride + a.xoff[field] : a.scratch;
    in.p[t] = (a.P && valid) ? a.P + (size_t)b * N : a.scratch + N;
    in.ex[t] = a.EX + (((size_t)bc * 3 + field) * 2 + chain) * exs;
  }
  return in;
}

// ---------------------------------------------------------------------------
// k_chain_local: one segment with zero entry state.
// ---------------------------------------------------------------------------
template <int NRS, int NB, int KD = kDepth>
__device__ __forceinline__ void local_nab2(const ChainArgs& a, int field, int start, int k0,
                                           int k1, const Inst<NB>& in,
                                           double (*xs)[NB][NRS + 64], double* exit_seg) {
  using S = CShape<NRS>;
  const int K = a.K, nr = a.nr;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int r = 16 * w + (lane >> 2), kq = lane & 3;
  const bool rv = r < nr;
  const int rr = rv ? r : nr - 1;
  // lanes of padding rows read the operator entries of row nr-1 (same cache
  // lines as the valid lanes, so no extra memory traffic); their results
  // only meet zero operator columns and are never stored
  const int lq = lane - 4 * (r - rr);
  const bool writer = kq == 0 && r < NRS;
  const int wp = writer ? vpos(r, S::CPL) : NRS + lane;
  const double dtf = a.dtf[field];
  // wave-uniform row bases, lane parts as byte offsets
  const double* lb = a.Lq[field] + w * S::NQ * 128;
  const unsigned lo = 16u * lq, go = 8u * rr, so = 8u * r;
  double2 blk[KD][S::NQ];
  double gv[KD][NB];
  auto fetch = [&](int k, double2 (&bq)[S::NQ], double (&g)[NB]) {
    const int j = start - 2 * (k < k1 ? k : k1 - 1);
    load_blk<S::NQ>(bq, lane_at(lb + (size_t)(K - 1 - j) * S::BLK, opaque(lo)));
    const unsigned gq = opaque(go);
#pragma unroll
    for (int t = 0; t < NB; ++t) g[t] = *lane_at(in.g[t] + (size_t)j * nr, gq);
  };
#pragma unroll
  for (int d = 0; d < KD; ++d) fetch(k0 + d, blk[d], gv[d]);
  drain_loads();
  double b[NB];
#pragma unroll
  for (int t = 0; t < NB; ++t) b[t] = 0.0;
  // one row of the chain; live is false only for the rows of the last group
  // that lie past the segment's end (uniform over the workgroup)
  auto row = [&](int k, int d, bool live) {
    const int j = start - 2 * (live ? k : k1 - 1);
    // NAB2 (Matrix_Operators.py:1061-1083): f_j = Linv (g_j - c_j b),
    // c_0 = 1/2; then b += 2 dt j f_j for row j-2.
    const double cj = (j == 0) ? 0.5 : 1.0;
#pragma unroll
    for (int t = 0; t < NB; ++t) xs[k & 1][t][wp] = gv[d][t] - cj * b[t];
    lds_barrier();
    const double alpha = 2.0 * dtf * j;
    double fr[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) fr[t] = pdot<S::NQ>(blk[d], &xs[k & 1][t][kq * S::CPL]);
#pragma unroll
    for (int t = 0; t < NB; ++t) fr[t] = quad_sum(fr[t]);
    if (live) {
      if (rv) {
        const unsigned sq = opaque(so);
#pragma unroll
        for (int t = 0; t < NB; ++t) *lane_at(in.x[t] + (size_t)j * nr, sq) = fr[t];
      }
#pragma unroll
      for (int t = 0; t < NB; ++t) b[t] = b[t] + alpha * fr[t];
    }
    fetch(k + KD, blk[d], gv[d]);
  };
  int kb = k0;
  for (; kb + KD <= k1; kb += KD) {
#pragma unroll
    for (int d = 0; d < KD; ++d) row(kb + d, d, true);
  }
  if (kb < k1) {
#pragma unroll
    for (int d = 0; d < KD; ++d) row(kb + d, d, kb + d < k1);
  }
  if (rv) {
#pragma unroll
    for (int t = 0; t < NB; ++t) *lane_at(in.ex[t] + (exit_seg - a.EX), so) = b[t];
  }
}

template <int NRS, int NB, int KD = kDepth>
__device__ __forceinline__ void local_a4(const ChainArgs& a, int start, int k0, int k1,
                                         const Inst<NB>& in, double (*xs)[NB][NRS + 64],
                                         double* exit_seg) {
  using S = CShape<NRS>;
  const int K = a.K, nr = a.nr;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int r = 16 * w + (lane >> 2), kq = lane & 3;
  const bool rv = r < nr;
  const int rr = rv ? r : nr - 1;
  // lanes of padding rows read the operator entries of row nr-1 (same cache
  // lines as the valid lanes, so no extra memory traffic); their results
  // only meet zero operator columns and are never stored
  const int lq = lane - 4 * (r - rr);
  const bool writer = kq == 0 && r < NRS;
  const int wp = writer ? vpos(r, S::CPL) : NRS + lane;
  const double dta = a.dtf[0];
  const double ir4 = a.vecs[V_IR4A4 * S::NRP + rr];
  const double ir2 = a.vecs[V_IR2A4 * S::NRP + rr];
  const double* lb = a.Lq[0] + w * S::NQ * 128;
  const double* mb = a.DLq + w * S::NQ * 128;
  const unsigned lo = 16u * lq, go = 8u * rr;
  double* sink = a.scratch + 2 * (size_t)K * nr + threadIdx.x;
  double2 blk[KD][S::NQ], mblk[KD][S::NQ];
  double gv[KD][NB];
  auto fetch = [&](int k, double2 (&bq)[S::NQ], double2 (&mq)[S::NQ], double (&g)[NB]) {
    const int j = start - 2 * (k < k1 ? k : k1 - 1);
    load_blk<S::NQ>(bq, lane_at(lb + (size_t)(K - 1 - j) * S::BLK, opaque(lo)));
    load_blk<S::NQ>(mq, lane_at(mb + (size_t)(K - 1 - j) * S::BLK, opaque(lo)));
    const unsigned gq = opaque(go);
#pragma unroll
    for (int t = 0; t < NB; ++t) g[t] = *lane_at(in.g[t] + (size_t)j * nr, gq);
  };
#pragma unroll
  for (int d = 0; d < KD; ++d) fetch(k0 + d, blk[d], mblk[d], gv[d]);
  drain_loads();
  double fe[NB], bfe[NB], dfe[NB];
#pragma unroll
  for (int t = 0; t < NB; ++t) fe[t] = bfe[t] = dfe[t] = 0.0;
  for (int kb = k0; kb < k1; kb += KD) {
#pragma unroll
    for (int d = 0; d < KD; ++d) {
      const int k = kb + d;
      const bool live = k < k1;
      const int row = start - 2 * (live ? k : k1 - 1);
      // A4 (Matrix_Operators.py:1141-1164), dt = Pr dt, j = row + 1:
      // b = dt bjt (L1 f_e + IR4 bf_e) - bjt IR2 f_e with L1 = D2 + bj IR4;
      // f = Linv (g + b); bf_e += bj f + bjt f_e; f_e += f.  f_e is also the
      // same-parity suffix sum P of the row.  D2 f_e is carried as a state
      // too: D2 f = (D2 Linv) (g + b) comes from the same right-hand side as
      // f, so a row needs one exchange of one vector.
      const double jp = (double)(row + 1);
      const double bj = -jp * (jp + 1.0);
      const double bjt = -2.0 * jp;
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        double* pdst = (live && rv) ? in.p[t] + (size_t)row * nr + r : sink;
        *pdst = fe[t];
        const double l1fe = dfe[t] + bj * (ir4 * fe[t]);
        const double bb = (dta * bjt) * (l1fe + ir4 * bfe[t]) - bjt * (ir2 * fe[t]);
        xs[k & 1][t][wp] = gv[d][t] + bb;
      }
      lds_barrier();
      double fr[NB], dr[NB];
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        fr[t] = pdot<S::NQ>(blk[d], &xs[k & 1][t][kq * S::CPL]);
        dr[t] = pdot<S::NQ>(mblk[d], &xs[k & 1][t][kq * S::CPL]);
      }
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        fr[t] = quad_sum(fr[t]);
        dr[t] = quad_sum(dr[t]);
      }
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        const double f = fr[t];
        const double df = dr[t];
        double* dst = (live && rv) ? in.x[t] + (size_t)row * nr + r : sink;
        *dst = f;
        const double nbfe = bfe[t] + bj * f + bjt * fe[t];
        bfe[t] = live ? nbfe : bfe[t];
        fe[t] = live ? fe[t] + f : fe[t];
        dfe[t] = live ? dfe[t] + df : dfe[t];
      }
      fetch(k + KD, blk[d], mblk[d], gv[d]);
    }
  }
  const size_t off = exit_seg - a.EX;
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    double* df = rv ? in.ex[t] + off + r : sink;
    double* db = rv ? in.ex[t] + off + S::NRP + r : sink;
    *df = fe[t];
    *db = bfe[t];
  }
}

// The A4 chain with two exchanges per row: D2 f_e comes from f_e itself,
// with D2 (the same block for every mode) held in registers, so a row
// streams one operator block instead of two.  Same recurrence and results
// as local_a4 up to summation order.
template <int NRS, int NB, int KD = kDepth>
__device__ __forceinline__ void local_a4x(const ChainArgs& a, int start, int k0, int k1,
                                          const Inst<NB>& in, double (*xs)[NB][NRS + 64],
                                          double* exit_seg) {
  using S = CShape<NRS>;
  const int K = a.K, nr = a.nr;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int r = 16 * w + (lane >> 2), kq = lane & 3;
  const bool rv = r < nr;
  const int rr = rv ? r : nr - 1;
  // lanes of padding rows read the operator entries of row nr-1 (same cache
  // lines as the valid lanes, so no extra memory traffic); their results
  // only meet zero operator columns and are never stored
  const int lq = lane - 4 * (r - rr);
  const bool writer = kq == 0 && r < NRS;
  const int wp = writer ? vpos(r, S::CPL) : NRS + lane;
  const double dta = a.dtf[0];
  const double ir4 = a.vecs[V_IR4A4 * S::NRP + rr];
  const double ir2 = a.vecs[V_IR2A4 * S::NRP + rr];
  // wave-uniform row bases, lane parts as byte offsets
  const unsigned lo = 16u * lq, go = 8u * rr, so = 8u * r;
  double2 d2[S::NQ];
  load_blk<S::NQ>(d2, lane_at(a.D2q + w * S::NQ * 128, lo));
  const double* lb = a.Lq[0] + w * S::NQ * 128;
  double2 blk[KD][S::NQ];
  double gv[KD][NB];
  auto fetch = [&](int k, double2 (&bq)[S::NQ], double (&g)[NB]) {
    const int j = start - 2 * (k < k1 ? k : k1 - 1);
    load_blk<S::NQ>(bq, lane_at(lb + (size_t)(K - 1 - j) * S::BLK, opaque(lo)));
    const unsigned gq = opaque(go);
#pragma unroll
    for (int t = 0; t < NB; ++t) g[t] = *lane_at(in.g[t] + (size_t)j * nr, gq);
  };
#pragma unroll
  for (int d = 0; d < KD; ++d) fetch(k0 + d, blk[d], gv[d]);
  drain_loads();
  double fe[NB], bfe[NB];
#pragma unroll
  for (int t = 0; t < NB; ++t) fe[t] = bfe[t] = 0.0;
  // one row of the chain; live is false only for the rows of the last group
  // that lie past the segment's end (uniform over the workgroup)
  auto step_row = [&](int k, int d, bool live) {
    const int row = start - 2 * (live ? k : k1 - 1);
    // A4 (Matrix_Operators.py:1141-1164), dt = Pr dt, j = row + 1:
    // b = dt bjt (L1 f_e + IR4 bf_e) - bjt IR2 f_e with L1 = D2 + bj IR4;
    // f = Linv (g + b); bf_e += bj f + bjt f_e; f_e += f.  f_e is also the
    // same-parity suffix sum P of the row.  Buffer 0 carries f_e, buffer 1
    // the right-hand side; the barrier after buffer 1 is written orders
    // every read of buffer 0 before the next row overwrites it.
    const double jp = (double)(row + 1);
    const double bj = -jp * (jp + 1.0);
    const double bjt = -2.0 * jp;
    if (live && rv) {
      const unsigned sq = opaque(so);
#pragma unroll
      for (int t = 0; t < NB; ++t) *lane_at(in.p[t] + (size_t)row * nr, sq) = fe[t];
    }
#pragma unroll
    for (int t = 0; t < NB; ++t) xs[0][t][wp] = fe[t];
    lds_barrier();
    // the instances' products in separate loops, so that their dependent
    // chains interleave instead of running one after the other
    double d2fe[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) d2fe[t] = pdot<S::NQ>(d2, &xs[0][t][kq * S::CPL]);
#pragma unroll
    for (int t = 0; t < NB; ++t) d2fe[t] = quad_sum(d2fe[t]);
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      const double l1fe = d2fe[t] + bj * (ir4 * fe[t]);
      const double bb = (dta * bjt) * (l1fe + ir4 * bfe[t]) - bjt * (ir2 * fe[t]);
      xs[1][t][wp] = gv[d][t] + bb;
    }
    lds_barrier();
    double fr[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) fr[t] = pdot<S::NQ>(blk[d], &xs[1][t][kq * S::CPL]);
#pragma unroll
    for (int t = 0; t < NB; ++t) fr[t] = quad_sum(fr[t]);
    if (live) {
      if (rv) {
        const unsigned sq = opaque(so);
#pragma unroll
        for (int t = 0; t < NB; ++t) *lane_at(in.x[t] + (size_t)row * nr, sq) = fr[t];
      }
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        const double f = fr[t];
        bfe[t] = bfe[t] + bj * f + bjt * fe[t];
        fe[t] = fe[t] + f;
      }
    }
    fetch(k + KD, blk[d], gv[d]);
  };
  int kb = k0;
  for (; kb + KD <= k1; kb += KD) {
#pragma unroll
    for (int d = 0; d < KD; ++d) step_row(kb + d, d, true);
  }
  if (kb < k1) {
#pragma unroll
    for (int d = 0; d < KD; ++d) step_row(kb + d, d, kb + d < k1);
  }
  const size_t off = exit_seg - a.EX;
  if (rv) {
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      *lane_at(in.ex[t] + off, so) = fe[t];
      *lane_at(in.ex[t] + off + S::NRP, so) = bfe[t];
    }
  }
}

template <int NRS, int NB>
__global__ __launch_bounds__(64 * CShape<NRS>::WR) void k_chain_local(ChainArgs a) {
  using S = CShape<NRS>;
  __shared__ __align__(16) double xs[2][NB][NRS + 64];
  const int field = blockIdx.x >> 1, chain = blockIdx.x & 1, seg = blockIdx.y;
  if (!((a.fields_mask >> field) & 1)) return;
  const int start = a.K - 1 - chain;
  const int n = start / 2 + 1;
  const int k0 = seg * a.m;
  if (k0 >= n) return;
  const int k1 = k0 + a.m < n ? k0 + a.m : n;
  const Inst<NB> in = instances<NB>(a, field, chain, blockIdx.z);
  // Eq_SYM keeps odd psi rows (A4 chain from K-1) and even T/S rows (NAB2
  // chain from K-2); the other chain is zero (Matrix_Operators.py:1054,1166).
  if (a.symmetric && (field == 0 ? chain == 1 : chain == 0)) {
    const int lane = threadIdx.x & 63;
    const int r = 16 * (threadIdx.x >> 6) + (lane >> 2);
    if ((lane & 3) == 0 && r < a.nr)
      for (int k = k0; k < k1; ++k) {
        const size_t o = (size_t)(start - 2 * k) * a.nr + r;
#pragma unroll
        for (int t = 0; t < NB; ++t) {
          in.x[t][o] = 0.0;
          if (field == 0) in.p[t][o] = 0.0;
        }
      }
    return;
  }
  double* exit_seg = a.EX + (size_t)seg * 2 * S::NRP;   // offset within an instance's (field, chain)
  if (field == 0)
    local_a4<NRS, NB>(a, start, k0, k1, in, xs, exit_seg);
  else
    local_nab2<NRS, NB>(a, field, start, k0, k1, in, xs, exit_seg);
}

// k_chain_local with two widths in one launch: the A4 chains take NBA
// instances per workgroup (X2: the two-exchange form, one streamed block per
// row), the NAB2 chains NBN.  The launch has ceil(B / min(NBA, NBN)) groups;
// workgroups past the end of the batch for their width return at once.
//   solve_instances 3: NBA 2, NBN 4 (one exchange per A4 row), KD 2
//   solve_instances 5: NBA 2, NBN 4, X2, KD 3
//   solve_instances 6: NBA 3, NBN 4, X2, KD 2
// KD is the number of rows of operator data in flight; it changes when data
// arrives, not the arithmetic.  Measured at N_fm = 512, N_r = 40, B = 256
// with KD 2: the full-batch solve takes 374 µs (3), 362 µs (5) and 303 µs
// (6); the A4 chains stream a third (5) and two ninths (6) of 3's bytes per
// instance.  Inside the four-stream step graph, width 5 with KD 3 is the
// fastest (567 µs per step against 613 µs for width 6 with KD 2; KD 3 made
// width 6 slower and KD 4 made width 5 slower).  Four A4 or eight NAB2
// instances per workgroup were slower (more per-row work on three
// wavefronts), as were three NAB2 instances or one A4 instance with KD 3.
// Work items of a k_chain_local_mixed launch, which is one dimensional and
// holds only live items: per (field, chain) type, nseg segments of
// ceil(B / width) instance groups (width NBA for the A4 chains, NBN for the
// T and S chains); types in order, then segments, then groups.  Workgroups
// go to the eight XCDs round-robin (workgroup w to XCD w % 8), so the list is
// cut into eight runs of consecutive items and run c is given to the
// workgroups of XCD c: each XCD's L2 then streams the operator blocks of one
// or two (type, segment)s.  In (type, segment, group) grid order four XCDs
// fetched every type's blocks (the solve read 1.7x its bytes).
struct MixedItem {
  int field, chain, seg, group;   // field -1: padding past the list
};

__host__ __device__ inline int mixed_types(int fields_mask) { return fields_mask == 1 ? 2 : 6; }

__host__ __device__ inline int mixed_item_count(int types, int nseg, int za, int zn) {
  return nseg * (2 * za + (types - 2) * zn);
}

__host__ __device__ inline int mixed_grid(int types, int nseg, int za, int zn) {
  return (mixed_item_count(types, nseg, za, zn) + 7) / 8 * 8;
}

__device__ __forceinline__ MixedItem mixed_item(int w, int types, int nseg, int za, int zn) {
  const int total = mixed_item_count(types, nseg, za, zn);
  const int i = (w % 8) * (mixed_grid(types, nseg, za, zn) / 8) + w / 8;
  MixedItem it{-1, 0, 0, 0};
  if (i >= total) return it;
  const int na = nseg * za;
  int type, r, z;
  if (i < 2 * na) {
    type = i / na;
    r = i - type * na;
    z = za;
  } else {
    const int nn = nseg * zn, j = i - 2 * na;
    type = 2 + j / nn;
    r = j - (type - 2) * nn;
    z = zn;
  }
  it.field = type >> 1;
  it.chain = type & 1;
  it.seg = r / z;
  it.group = r - it.seg * z;
  return it;
}

template <int NRS, int NBA, int NBN, bool X2, int KD>
__global__ __launch_bounds__(64 * CShape<NRS>::WR) void k_chain_local_mixed(ChainArgs a) {
  using S = CShape<NRS>;
  constexpr int NBX = NBA > NBN ? NBA : NBN;
  __shared__ __align__(16) double xs[2][NBX][NRS + 64];
  const MixedItem it = mixed_item(blockIdx.x, mixed_types(a.fields_mask), a.nseg,
                                  (a.B + NBA - 1) / NBA, (a.B + NBN - 1) / NBN);
  const int field = it.field, chain = it.chain, seg = it.seg;
  if (field < 0 || !((a.fields_mask >> field) & 1)) return;
  const int start = a.K - 1 - chain;
  const int n = start / 2 + 1;
  const int k0 = seg * a.m;
  if (k0 >= n) return;
  const int k1 = k0 + a.m < n ? k0 + a.m : n;
  double* exit_seg = a.EX + (size_t)seg * 2 * S::NRP;
  const bool sym_zero = a.symmetric && (field == 0 ? chain == 1 : chain == 0);
  if (field == 0) {
    const Inst<NBA> in = instances<NBA>(a, field, chain, it.group);
    if (sym_zero) {   // Eq_SYM: this chain is zero (Matrix_Operators.py:1166)
      const int lane = threadIdx.x & 63;
      const int r = 16 * (threadIdx.x >> 6) + (lane >> 2);
      if ((lane & 3) == 0 && r < a.nr)
        for (int k = k0; k < k1; ++k) {
          const size_t o = (size_t)(start - 2 * k) * a.nr + r;
#pragma unroll
          for (int t = 0; t < NBA; ++t) {
            in.x[t][o] = 0.0;
            in.p[t][o] = 0.0;
          }
        }
      return;
    }
    auto* xa = reinterpret_cast<double (*)[NBA][NRS + 64]>(xs);
    if constexpr (X2)
      local_a4x<NRS, NBA, KD>(a, start, k0, k1, in, xa, exit_seg);
    else
      local_a4<NRS, NBA, KD>(a, start, k0, k1, in, xa, exit_seg);
    return;
  }
  const Inst<NBN> in = instances<NBN>(a, field, chain, it.group);
  if (sym_zero) {   // Eq_SYM: this chain is zero (Matrix_Operators.py:1054)
    const int lane = threadIdx.x & 63;
    const int r = 16 * (threadIdx.x >> 6) + (lane >> 2);
    if ((lane & 3) == 0 && r < a.nr)
      for (int k = k0; k < k1; ++k) {
        const size_t o = (size_t)(start - 2 * k) * a.nr + r;
#pragma unroll
        for (int t = 0; t < NBN; ++t) in.x[t][o] = 0.0;
      }
    return;
  }
  local_nab2<NRS, NBN, KD>(a, field, start, k0, k1, in,
                                    reinterpret_cast<double (*)[NBN][NRS + 64]>(xs), exit_seg);
}

// ---------------------------------------------------------------------------
// k_chain_join: entry state of a segment, then the row corrections.
// ---------------------------------------------------------------------------
template <int NRS, int NB>
__device__ __forceinline__ void join_nab2(const ChainArgs& a, int field, int chain, int start,
                                          int seg, int k0, int k1, const Inst<NB>& in,
                                          double (*xs)[NB][2 * NRS + 64]) {
  using S = CShape<NRS>;
  const int K = a.K, nr = a.nr;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int r = 16 * w + (lane >> 2), kq = lane & 3;
  const bool rv = r < nr;
  const int rr = rv ? r : nr - 1;
  // lanes of padding rows read the operator entries of row nr-1 (same cache
  // lines as the valid lanes, so no extra memory traffic); their results
  // only meet zero operator columns and are never stored
  const int lq = lane - 4 * (r - rr);
  const bool writer = kq == 0 && r < NRS;
  const int wp = writer ? vpos(r, S::CPL) : NRS + lane;
  double* sink = a.scratch + 2 * (size_t)K * nr + threadIdx.x;
  // entry state: s_0 = 0, s_{u+1} = ex_u + E_u s_u for u < seg
  const double* eb = a.Eq[field] + (size_t)chain * a.nseg * S::BLK + w * S::NQ * 128 + 2 * lq;
  double2 e[kDepthE][S::NQ];
  double exv[kDepthE][NB];
  auto fetch_e = [&](int u, double2 (&eq)[S::NQ], double (&x)[NB]) {
    const int uc = u < seg ? u : seg - 1;
    load_blk<S::NQ>(eq, eb + (size_t)uc * S::BLK);
#pragma unroll
    for (int t = 0; t < NB; ++t) x[t] = in.ex[t][(size_t)uc * 2 * S::NRP + rr];
  };
#pragma unroll
  for (int d = 0; d < kDepthE; ++d) fetch_e(d, e[d], exv[d]);
  drain_loads();
  double sv[NB];
#pragma unroll
  for (int t = 0; t < NB; ++t) sv[t] = 0.0;
  for (int ub = 0; ub < seg; ub += kDepthE) {
#pragma unroll
    for (int d = 0; d < kDepthE; ++d) {
      const bool live = ub + d < seg;
#pragma unroll
      for (int t = 0; t < NB; ++t)
        xs[d & 1][t][wp] = sv[t];
      lds_barrier();
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        const double nv = exv[d][t] + quad_sum(pdot<S::NQ>(e[d], &xs[d & 1][t][kq * S::CPL]));
        sv[t] = live ? nv : sv[t];
      }
      fetch_e(ub + d + kDepthE, e[d], exv[d]);
    }
  }
  // broadcast the entry state once (buffer 0: the prefix loop ran an even
  // number of exchanges)
#pragma unroll
  for (int t = 0; t < NB; ++t)
    xs[0][t][wp] = sv[t];
  lds_barrier();
  double xl[NB][2 * S::NQ];
#pragma unroll
  for (int t = 0; t < NB; ++t)
#pragma unroll
    for (int q = 0; q < 2 * S::NQ; ++q) xl[t][q] = xs[0][t][kq * S::CPL + q];
  // corrections f_row += Q_row s (independent rows)
  const double* qb = a.Qq[field] + w * S::NQ * 128 + 2 * lq;
  double2 qv[kDepth][S::NQ];
  double fl[kDepth][NB];
  auto fetch_q = [&](int k, double2 (&qq)[S::NQ], double (&f)[NB]) {
    const int j = start - 2 * (k < k1 ? k : k1 - 1);
    load_blk<S::NQ>(qq, qb + (size_t)(K - 1 - j) * S::BLK);
#pragma unroll
    for (int t = 0; t < NB; ++t) f[t] = in.x[t][(size_t)j * nr + rr];
  };
#pragma unroll
  for (int d = 0; d < kDepth; ++d) fetch_q(k0 + d, qv[d], fl[d]);
  drain_loads();
  for (int kb = k0; kb < k1; kb += kDepth) {
#pragma unroll
    for (int d = 0; d < kDepth; ++d) {
      const int k = kb + d;
      const bool live = k < k1;
      const int j = start - 2 * (live ? k : k1 - 1);
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        const double f = fl[d][t] + quad_sum(pdot_reg<S::NQ>(qv[d], xl[t]));
        double* dst = (live && rv) ? in.x[t] + (size_t)j * nr + r : sink;
        *dst = f;
      }
      fetch_q(k + kDepth, qv[d], fl[d]);
    }
  }
}

template <int NRS, int NB>
__device__ __forceinline__ void join_a4(const ChainArgs& a, int chain, int start, int seg,
                                        int k0, int k1, const Inst<NB>& in,
                                        double (*xs)[NB][2 * NRS + 64]) {
  using S = CShape<NRS>;
  constexpr int NQ2 = 2 * S::NQ;          // loads per lane of an NRP x 2NRS matrix
  constexpr int CPL2 = 2 * S::CPL;        // entries of a 2NRS vector per lane
  constexpr int BLK2 = 2 * S::BLK;        // doubles per packed NRP x 2NRS matrix
  const int K = a.K, nr = a.nr;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int r = 16 * w + (lane >> 2), kq = lane & 3;
  const bool rv = r < nr;
  const int rr = rv ? r : nr - 1;
  // lanes of padding rows read the operator entries of row nr-1 (same cache
  // lines as the valid lanes, so no extra memory traffic); their results
  // only meet zero operator columns and are never stored
  const int lq = lane - 4 * (r - rr);
  const bool writer = kq == 0 && r < NRS;
  const int wpf = writer ? vpos(r, CPL2) : 2 * NRS + lane;
  const int wpb = writer ? vpos(NRS + r, CPL2) : 2 * NRS + lane;
  double* sink = a.scratch + 2 * (size_t)K * nr + threadIdx.x;
  // state (f_e, bf_e): column c < NRS is f_e[c], column NRS + c is bf_e[c];
  // E is stored as its f_e rows then its bf_e rows.
  const double* eb = a.Eq[0] + (size_t)chain * a.nseg * 2 * BLK2 + w * NQ2 * 128 + 2 * lq;
  double2 ef[kDepthE][NQ2], ebv[kDepthE][NQ2];
  double exf[kDepthE][NB], exb[kDepthE][NB];
  auto fetch_e = [&](int u, double2 (&qf)[NQ2], double2 (&qb)[NQ2], double (&xf)[NB],
                     double (&xb)[NB]) {
    const int uc = u < seg ? u : seg - 1;
    load_blk<NQ2>(qf, eb + (size_t)uc * 2 * BLK2);
    load_blk<NQ2>(qb, eb + (size_t)uc * 2 * BLK2 + BLK2);
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      xf[t] = in.ex[t][(size_t)uc * 2 * S::NRP + rr];
      xb[t] = in.ex[t][(size_t)uc * 2 * S::NRP + S::NRP + rr];
    }
  };
#pragma unroll
  for (int d = 0; d < kDepthE; ++d) fetch_e(d, ef[d], ebv[d], exf[d], exb[d]);
  drain_loads();
  double sf[NB], sb[NB];
#pragma unroll
  for (int t = 0; t < NB; ++t) sf[t] = sb[t] = 0.0;
  for (int ub = 0; ub < seg; ub += kDepthE) {
#pragma unroll
    for (int d = 0; d < kDepthE; ++d) {
      const bool live = ub + d < seg;
#pragma unroll
      for (int t = 0; t < NB; ++t)
      {
        // lanes without an entry write their private tail slot (no branch)
        xs[d & 1][t][wpf] = sf[t];
        xs[d & 1][t][wpb] = sb[t];
      }
      lds_barrier();
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        const double* xv = &xs[d & 1][t][kq * CPL2];
        const double nf = exf[d][t] + quad_sum(pdot<NQ2>(ef[d], xv));
        const double nb = exb[d][t] + quad_sum(pdot<NQ2>(ebv[d], xv));
        sf[t] = live ? nf : sf[t];
        sb[t] = live ? nb : sb[t];
      }
      fetch_e(ub + d + kDepthE, ef[d], ebv[d], exf[d], exb[d]);
    }
  }
#pragma unroll
  for (int t = 0; t < NB; ++t)
  {
    xs[0][t][wpf] = sf[t];
    xs[0][t][wpb] = sb[t];
  }
  lds_barrier();
  double xl[NB][CPL2];
#pragma unroll
  for (int t = 0; t < NB; ++t)
#pragma unroll
    for (int q = 0; q < CPL2; ++q) xl[t][q] = xs[0][t][kq * CPL2 + q];
  // corrections f_row += Q_row s, and P_row = f_e of the row
  const double* qb = a.Qq[0] + w * NQ2 * 128 + 2 * lq;
  double2 qv[kDepth][NQ2];
  double fl[kDepth][NB];
  auto fetch_q = [&](int k, double2 (&qq)[NQ2], double (&f)[NB]) {
    const int j = start - 2 * (k < k1 ? k : k1 - 1);
    load_blk<NQ2>(qq, qb + (size_t)(K - 1 - j) * BLK2);
#pragma unroll
    for (int t = 0; t < NB; ++t) f[t] = in.x[t][(size_t)j * nr + rr];
  };
#pragma unroll
  for (int d = 0; d < kDepth; ++d) fetch_q(k0 + d, qv[d], fl[d]);
  drain_loads();
  double fe[NB];
#pragma unroll
  for (int t = 0; t < NB; ++t) fe[t] = sf[t];
  for (int kb = k0; kb < k1; kb += kDepth) {
#pragma unroll
    for (int d = 0; d < kDepth; ++d) {
      const int k = kb + d;
      const bool live = k < k1;
      const int j = start - 2 * (live ? k : k1 - 1);
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        const double f = fl[d][t] + quad_sum(pdot_reg<NQ2>(qv[d], xl[t]));
        double* dst = (live && rv) ? in.x[t] + (size_t)j * nr + r : sink;
        double* pdst = (live && rv) ? in.p[t] + (size_t)j * nr + r : sink;
        *dst = f;
        *pdst = fe[t];
        fe[t] = live ? fe[t] + f : fe[t];
      }
      fetch_q(k + kDepth, qv[d], fl[d]);
    }
  }
}

template <int NRS, int NB>
__global__ __launch_bounds__(64 * CShape<NRS>::WR) void k_chain_join(ChainArgs a) {
  __shared__ __align__(16) double xs[2][NB][2 * NRS + 64];
  // segments 1 .. nseg-1 (segment 0 starts from a zero state: nothing to join)
  const int field = blockIdx.x >> 1, chain = blockIdx.x & 1, seg = blockIdx.y + 1;
  if (!((a.fields_mask >> field) & 1)) return;
  if (a.symmetric && (field == 0 ? chain == 1 : chain == 0)) return;
  const int start = a.K - 1 - chain;
  const int n = start / 2 + 1;
  const int k0 = seg * a.m;
  if (k0 >= n) return;
  const int k1 = k0 + a.m < n ? k0 + a.m : n;
  const Inst<NB> in = instances<NB>(a, field, chain, blockIdx.z);
  if (field == 0)
    join_a4<NRS, NB>(a, chain, start, seg, k0, k1, in, xs);
  else
    join_nab2<NRS, NB>(a, field, chain, start, seg, k0, k1, in, xs);
}

// ---------------------------------------------------------------------------
// k_chain_mfma: the local (segment) chain solve for 16 instances per
// workgroup on MFMA.  Wavefront w owns block rows 16w..16w+15 as one M tile
// of __builtin_amdgcn_mfma_f64_16x16x4f64; the 16 instances are the N columns.
// A fragments: lane l holds block[16w + l%16][4ks + l/16] (MFMA-packed, one
// 16-byte load per two k steps); B fragments: the right-hand sides, exchanged
// through LDS as [row][instance]; output: lane l, register v holds row
// 16w + l/16 + 4v of instance l%16, which is also where the running state of
// the chain lives.  Results equal the quad kernel's up to summation order.
// ---------------------------------------------------------------------------
typedef double v4d_t __attribute__((ext_vector_type(4)));

template <int NRS>
__device__ __forceinline__ void mfma_block(const double2 (&a)[NRS / 8], const double* xs,
                                           int lane, v4d_t& acc) {
  // xs: [NRS][16] right-hand sides; B fragment of k step ks: row 4ks + lane/16
#pragma unroll
  for (int kp = 0; kp < NRS / 8; ++kp) {
    const double b0 = xs[(8 * kp + (lane >> 4)) * 16 + (lane & 15)];
    const double b1 = xs[(8 * kp + 4 + (lane >> 4)) * 16 + (lane & 15)];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kp].x, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kp].y, b1, acc, 0, 0, 0);
  }
}

// Two products of one right-hand side: acc += A xs, accd += M xs (each B
// fragment read once; the two MFMA chains are independent).
template <int NRS>
__device__ __forceinline__ void mfma_block2(const double2 (&a)[NRS / 8], const double2 (&m)[NRS / 8],
                                            const double* xs, int lane, v4d_t& acc, v4d_t& accd) {
#pragma unroll
  for (int kp = 0; kp < NRS / 8; ++kp) {
    const double b0 = xs[(8 * kp + (lane >> 4)) * 16 + (lane & 15)];
    const double b1 = xs[(8 * kp + 4 + (lane >> 4)) * 16 + (lane & 15)];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kp].x, b0, acc, 0, 0, 0);
    accd = __builtin_amdgcn_mfma_f64_16x16x4f64(m[kp].x, b0, accd, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kp].y, b1, acc, 0, 0, 0);
    accd = __builtin_amdgcn_mfma_f64_16x16x4f64(m[kp].y, b1, accd, 0, 0, 0);
  }
}

template <int NRS>
__global__ __launch_bounds__(64 * CShape<NRS>::WR) void k_chain_mfma(ChainArgs a) {
  using S = CShape<NRS>;
  constexpr int NKP = NRS / 8;                 // 16-byte A loads per lane per block
  constexpr int MBLK = S::NRP * NRS;           // doubles per MFMA-packed block
  __shared__ __align__(16) double xs[2][NRS * 16 + 64];
  const int field = blockIdx.x >> 1, chain = blockIdx.x & 1, seg = blockIdx.y;
  if (!((a.fields_mask >> field) & 1)) return;
  const int K = a.K, nr = a.nr;
  const int start = K - 1 - chain;
  const int n = start / 2 + 1;
  const int k0 = seg * a.m;
  if (k0 >= n) return;
  const int k1 = k0 + a.m < n ? k0 + a.m : n;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int c = lane & 15;                     // instance column
  const int b = blockIdx.z * 16 + c;
  const bool valid = b < a.B;
  const int bc = valid ? b : a.B - 1;
  const size_t N = (size_t)K * nr;
  const double* gp = a.G + (size_t)bc * a.gstride + a.goff[field];
  double* xp = valid ? a.X + (size_t)b * a.xstride + a.xoff[field] : a.scratch;
  double* pp = (a.P && valid) ? a.P + (size_t)b * N : a.scratch + N;
  double* exb = a.EX + ((((size_t)(valid ? b : bc) * 3 + field) * 2 + chain) * a.nseg + seg) *
                           2 * S::NRP;
  double* sink = a.scratch + 2 * N + threadIdx.x;
  int rv[4], rrv[4], wpos[4];
  bool okr[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    rv[v] = 16 * w + (lane >> 4) + 4 * v;
    okr[v] = rv[v] < nr;
    rrv[v] = okr[v] ? rv[v] : nr - 1;
    wpos[v] = rv[v] < NRS ? rv[v] * 16 + c : NRS * 16 + lane;   // tail slot for rows >= NRS
  }
  const bool sym_skip = a.symmetric && (field == 0 ? chain == 1 : chain == 0);
  if (sym_skip) {   // Eq_SYM: this chain is zero (Matrix_Operators.py:1054,1166)
    for (int k = k0; k < k1; ++k) {
      const int j = start - 2 * k;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        *((okr[v] && valid) ? xp + (size_t)j * nr + rv[v] : sink) = 0.0;
        if (field == 0) *((okr[v] && valid) ? pp + (size_t)j * nr + rv[v] : sink) = 0.0;
      }
    }
    return;
  }
  const double* lb = a.Lm[field] + (size_t)w * NKP * 128 + 2 * lane;
  double2 blk[kDepthM][NKP];
  double gv[kDepthM][4];
  auto fetch = [&](int k, double2 (&bq)[NKP], double (&g)[4]) {
    const int j = start - 2 * (k < k1 ? k : k1 - 1);
    const double* p = lb + (size_t)(K - 1 - j) * MBLK;
#pragma unroll
    for (int q = 0; q < NKP; ++q) bq[q] = *reinterpret_cast<const double2*>(p + q * 128);
#pragma unroll
    for (int v = 0; v < 4; ++v) g[v] = gp[(size_t)j * nr + rrv[v]];
  };
#pragma unroll
  for (int d = 0; d < kDepthM; ++d) fetch(k0 + d, blk[d], gv[d]);
  if (field != 0) {
    // NAB2 (Matrix_Operators.py:1061-1083)
    drain_loads();
    const double dtf = a.dtf[field];
    double bs[4] = {0.0, 0.0, 0.0, 0.0};
    for (int kb = k0; kb < k1; kb += kDepthM) {
#pragma unroll
      for (int d = 0; d < kDepthM; ++d) {
        const int k = kb + d;
        const bool live = k < k1;
        const int j = start - 2 * (live ? k : k1 - 1);
        const double cj = (j == 0) ? 0.5 : 1.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) xs[d & 1][wpos[v]] = gv[d][v] - cj * bs[v];
        lds_barrier();
        v4d_t acc = {0.0, 0.0, 0.0, 0.0};
        mfma_block<NRS>(blk[d], xs[d & 1], lane, acc);
        const double alpha = 2.0 * dtf * j;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          *((live && okr[v] && valid) ? xp + (size_t)j * nr + rv[v] : sink) = acc[v];
          bs[v] = live ? bs[v] + alpha * acc[v] : bs[v];
        }
        fetch(k + kDepthM, blk[d], gv[d]);
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) *((okr[v] && valid) ? exb + rv[v] : sink) = bs[v];
    return;
  }
  // A4 (Matrix_Operators.py:1141-1164), dt = Pr dt; D2 f_e carried as a
  // state (D2 f = (D2 Linv_j) rhs from the same right-hand side as f)
  const double dta = a.dtf[0];
  const double* mbase = a.DLm + (size_t)w * NKP * 128 + 2 * lane;
  double2 mfr[kDepthM][NKP];
  auto fetchm = [&](int k, double2 (&mq)[NKP]) {
    const int j = start - 2 * (k < k1 ? k : k1 - 1);
    const double* p = mbase + (size_t)(K - 1 - j) * MBLK;
#pragma unroll
    for (int q = 0; q < NKP; ++q) mq[q] = *reinterpret_cast<const double2*>(p + q * 128);
  };
#pragma unroll
  for (int d = 0; d < kDepthM; ++d) fetchm(k0 + d, mfr[d]);
  double ir4[4], ir2[4], fe[4], bfe[4], dfe[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    ir4[v] = a.vecs[V_IR4A4 * S::NRP + rrv[v]];
    ir2[v] = a.vecs[V_IR2A4 * S::NRP + rrv[v]];
    fe[v] = bfe[v] = dfe[v] = 0.0;
  }
  drain_loads();
  for (int kb = k0; kb < k1; kb += kDepthM) {
#pragma unroll
    for (int d = 0; d < kDepthM; ++d) {
      const int k = kb + d;
      const bool live = k < k1;
      const int row = start - 2 * (live ? k : k1 - 1);
      const double jp = (double)(row + 1);
      const double bj = -jp * (jp + 1.0);
      const double bjt = -2.0 * jp;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        *((live && okr[v] && valid) ? pp + (size_t)row * nr + rv[v] : sink) = fe[v];
        const double l1fe = dfe[v] + bj * (ir4[v] * fe[v]);
        const double bb = (dta * bjt) * (l1fe + ir4[v] * bfe[v]) - bjt * (ir2[v] * fe[v]);
        xs[d & 1][wpos[v]] = gv[d][v] + bb;
      }
      lds_barrier();
      v4d_t acc = {0.0, 0.0, 0.0, 0.0}, accd = {0.0, 0.0, 0.0, 0.0};
      mfma_block2<NRS>(blk[d], mfr[d], xs[d & 1], lane, acc, accd);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        *((live && okr[v] && valid) ? xp + (size_t)row * nr + rv[v] : sink) = acc[v];
        const double nbfe = bfe[v] + bj * acc[v] + bjt * fe[v];
        bfe[v] = live ? nbfe : bfe[v];
        fe[v] = live ? fe[v] + acc[v] : fe[v];
        dfe[v] = live ? dfe[v] + accd[v] : dfe[v];
      }
      fetch(k + kDepthM, blk[d], gv[d]);
      fetchm(k + kDepthM, mfr[d]);
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    *((okr[v] && valid) ? exb + rv[v] : sink) = fe[v];
    *((okr[v] && valid) ? exb + S::NRP + rv[v] : sink) = bfe[v];
  }
}

// Chain workgroups along x: (field, chain) pairs; two when only the A4
// solve runs (the NAB2 fields in the eigenbasis or not asked for).
static inline unsigned chain_grid_x(const ChainArgs& a) { return a.fields_mask == 1 ? 2u : 6u; }

// ---------------------------------------------------------------------------
// NAB2 solves in the eigenbasis (Plan::nab2_modal).  The block of row j is
// Linv_j = (P + h j(j+1) I)^-1 with P = V diag(lam) V^-1, so in the variables
// gt = V^-1 g, ft = V^-1 f, bt = V^-1 b the chain of NAB2_BSub_TSTEP_V2
// (Matrix_Operators.py:1061-1083)
//   f_j = Linv_j (g_j - c_j b),  then b += 2 h j f_j for row j - 2
// is nr independent scalar recurrences (c_j = 1/2 at j = 0, else 1):
//   ft_j = (gt_j - c_j bt) q_j,  q_j = 1 / (lam + h j (j+1)),
//   bt  <- bt + 2 h j ft_j = (lam + h j (j-1)) q_j bt + 2 h j q_j gt_j.
// One workgroup per (instance, field, chain) takes the chain in chunks of 64
// consecutive rows from the top.  Per chunk: the rows go to LDS (the next
// chunk's rows are loaded meanwhile); gt = g V^-T as four 16-row tiles times
// the wavefront's 16-column tile of V^-T on f64 MFMA (16x16x4: A = rows from
// LDS, B = V^-1 entries held in registers for the whole chain); the
// recurrences run on wavefront 0, one component per lane, carrying bt from
// chunk to chunk in registers; f = ft V^T the same way; the rows go back to
// memory.
// ---------------------------------------------------------------------------
struct ModalArgs {
  const double* G;
  double* X;
  long long gstride, xstride;
  long long goff[3], xoff[3];
  int K, nr, symmetric, field0;
  double h[3];
};

typedef double v4d_m __attribute__((ext_vector_type(4)));

// out[row][16w + l%16] = sum_e buf[row][e] M[16w + l%16][e] for the 64 rows
// of the chunk (lane l, register v of tile t: row 16t + l/16 + 4v), with
// bm[s] = M[16w + l%16][4s + l/16].  All reads of buf happen here; the
// caller writes the results after a barrier.
template <int NRS, int PITCH>
__device__ __forceinline__ void modal_tiles(const double* buf, const double (&bm)[NRS / 4],
                                            int lane, v4d_m (&acc)[4]) {
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = (v4d_m){0.0, 0.0, 0.0, 0.0};
  const double* arow = buf + (lane & 15) * PITCH + (lane >> 4);
#pragma unroll
  for (int s = 0; s < NRS / 4; ++s)
#pragma unroll
    for (int t = 0; t < 4; ++t)
      acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(arow[16 * t * PITCH + 4 * s], bm[s], acc[t],
                                                    0, 0, 0);
}

template <int NRS, int PITCH>
__device__ __forceinline__ void modal_store(double* buf, const v4d_m (&acc)[4], int col, int lane) {
  if (col >= NRS) return;   // columns past NRS do not exist in buf (and are zero)
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) buf[(16 * t + (lane >> 4) + 4 * v) * PITCH + col] = acc[t][v];
}

template <int NRS, int WR>
__global__ __launch_bounds__(64 * WR) void k_nab2_modal(ModalArgs a, const double* __restrict__ m1,
                                                        const double* __restrict__ m2) {
  constexpr int CH = 64;                  // rows per chunk
  constexpr int PITCH = NRS + 1;          // odd: column accesses spread over the banks
  constexpr int RPW = (CH + WR - 1) / WR; // rows of a chunk each wavefront loads
  __shared__ double buf[CH * PITCH];
  const int field = a.field0 + (int)(blockIdx.x >> 1), chain = blockIdx.x & 1;
  const int K = a.K, nr = a.nr;
  const int start = K - 1 - chain, n = start / 2 + 1;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const double* __restrict__ g = a.G + (size_t)blockIdx.y * a.gstride + a.goff[field];
  double* __restrict__ x = a.X + (size_t)blockIdx.y * a.xstride + a.xoff[field];
  if (a.symmetric && chain == 0) {   // Eq_SYM: odd T/S rows are zero (Matrix_Operators.py:1054)
    for (int i = w; i < n; i += WR)
      if (lane < nr) x[(size_t)(start - 2 * i) * nr + lane] = 0.0;
    return;
  }
  const double* __restrict__ md = field == 1 ? m1 : m2;
  const double h = a.h[field];
  const double lamk = md[lane < NRS ? lane : 0];
  // B operands of this wavefront's column tile: V^-1 [NRS][NRS] and V [nr][NRS]
  const int col = 16 * w + (lane & 15);
  double bvi[NRS / 4], bv[NRS / 4];
#pragma unroll
  for (int s = 0; s < NRS / 4; ++s) {
    const int e = 4 * s + (lane >> 4);
    bvi[s] = col < NRS ? md[NRS + (size_t)nr * NRS + (size_t)col * NRS + e] : 0.0;
    bv[s] = col < nr ? md[NRS + (size_t)col * NRS + e] : 0.0;
  }
  double pre[RPW];
  auto fetch = [&](int c0) {
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
      const int i = w + WR * u;
      const bool ok = i < CH && c0 + i < n && lane < nr;
      pre[u] = ok ? g[(size_t)(start - 2 * (c0 + i)) * nr + lane] : 0.0;
    }
  };
  fetch(0);
  double bt = 0.0;
  v4d_m acc[4];
  for (int c0 = 0; c0 < n; c0 += CH) {
    const int cn = n - c0 < CH ? n - c0 : CH;
#pragma unroll
    for (int u = 0; u < RPW; ++u) {   // padding rows and columns are zero
      const int i = w + WR * u;
      if (i < CH && lane < NRS) buf[i * PITCH + lane] = pre[u];
    }
    __syncthreads();
    if (c0 + CH < n) fetch(c0 + CH);
    modal_tiles<NRS, PITCH>(buf, bvi, lane, acc);
    __syncthreads();
    modal_store<NRS, PITCH>(buf, acc, col, lane);
    __syncthreads();
    if (w == 0 && lane < nr) {
      // eight rows at a time: their inputs and q are ready before the
      // recurrence reaches them (rows past the chunk are zero and skipped)
      for (int i0 = 0; i0 < cn; i0 += 8) {
        double gt[8], q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int j = start - 2 * (c0 + i0 + u);
          gt[u] = buf[(i0 + u) * PITCH + lane];
          q[u] = 1.0 / (lamk + (h * j) * (j + 1));
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if (i0 + u < cn) {
            const int j = start - 2 * (c0 + i0 + u);
            const double hj = h * j;
            const double cj = j == 0 ? 0.5 : 1.0;
            buf[(i0 + u) * PITCH + lane] = (gt[u] - cj * bt) * q[u];
            bt = fma((lamk + hj * (j - 1)) * q[u], bt, (2.0 * hj * q[u]) * gt[u]);
          }
        }
      }
    }
    __syncthreads();
    modal_tiles<NRS, PITCH>(buf, bv, lane, acc);
    __syncthreads();
    modal_store<NRS, PITCH>(buf, acc, col, lane);
    __syncthreads();
    for (int i = w; i < cn; i += WR)
      if (lane < nr) x[(size_t)(start - 2 * (c0 + i)) * nr + lane] = buf[i * PITCH + lane];
    __syncthreads();
  }
}

template <int NRS>
static int launch_modal_t(const ModalArgs& a, int nf, int B, const Plan& p, hipStream_t s) {
  using S = CShape<NRS>;
  const dim3 grid(2 * nf, B), block(64 * S::WR);
  hipLaunchKernelGGL((k_nab2_modal<NRS, S::WR>), grid, block, 0, s, a, p.nab2_modes[1],
                     p.nab2_modes[2]);
  return hip_check(hipGetLastError(), "k_nab2_modal");
}

static int dispatch_modal(const ModalArgs& a, int nf, int B, const Plan& p, hipStream_t s) {
  switch (p.NRS) {
    case 8: return launch_modal_t<8>(a, nf, B, p, s);
    case 16: return launch_modal_t<16>(a, nf, B, p, s);
    case 24: return launch_modal_t<24>(a, nf, B, p, s);
    case 32: return launch_modal_t<32>(a, nf, B, p, s);
    case 40: return launch_modal_t<40>(a, nf, B, p, s);
    case 48: return launch_modal_t<48>(a, nf, B, p, s);
    case 56: return launch_modal_t<56>(a, nf, B, p, s);
    case 64: return launch_modal_t<64>(a, nf, B, p, s);
    default: return fail(SDDC_ERR_UNSUPPORTED, "unsupported radial size");
  }
}

template <int NRS>
static int launch_chain_mfma_t(const ChainArgs& a, int B, hipStream_t s) {
  using S = CShape<NRS>;
  const dim3 grid(chain_grid_x(a), a.nseg, (B + 15) / 16), block(64 * S::WR);
  hipLaunchKernelGGL((k_chain_mfma<NRS>), grid, block, 0, s, a);
  if (int rc = hip_check(hipGetLastError(), "k_chain_mfma")) return rc;
  if (a.nseg > 1) {   // joins in the quad layout, two instances per workgroup
    const dim3 gj(chain_grid_x(a), a.nseg - 1, (B + 1) / 2);
    hipLaunchKernelGGL((k_chain_join<NRS, 2>), gj, block, 0, s, a);
    if (int rc = hip_check(hipGetLastError(), "k_chain_join")) return rc;
  }
  return SDDC_OK;
}

template <int NRS, int NB>
static int launch_chain_t(const ChainArgs& a, int B, hipStream_t s) {
  using S = CShape<NRS>;
  const dim3 grid(chain_grid_x(a), a.nseg, (B + NB - 1) / NB), block(64 * S::WR);
  hipLaunchKernelGGL((k_chain_local<NRS, NB>), grid, block, 0, s, a);
  if (int rc = hip_check(hipGetLastError(), "k_chain_local")) return rc;
  if (a.nseg > 1) {
    // four instances per join workgroup need 472 VGPRs at NRS = 40 (one
    // wavefront per SIMD) and spill to scratch from NRS = 48 on; there the
    // joins take two (the same arithmetic per instance).  Measured on the
    // library's default width 4 at 64 instances, N_r = 40: the solve took
    // 111 against 121 µs at N_fm = 256 and 197 against 201 µs at 512.
    constexpr int NBJ = (NB > 2 && NRS >= 40) ? 2 : NB;
    const dim3 gj(chain_grid_x(a), a.nseg - 1, (B + NBJ - 1) / NBJ);
    hipLaunchKernelGGL((k_chain_join<NRS, NBJ>), gj, block, 0, s, a);
    if (int rc = hip_check(hipGetLastError(), "k_chain_join")) return rc;
  }
  return SDDC_OK;
}

template <int NRS, int NBA, int NBN, bool X2, int KD>
static int launch_chain_mixed_t(const ChainArgs& a, int B, hipStream_t s) {
  using S = CShape<NRS>;
  const dim3 grid(mixed_grid(mixed_types(a.fields_mask), a.nseg, (B + NBA - 1) / NBA,
                             (B + NBN - 1) / NBN)),
      block(64 * S::WR);
  hipLaunchKernelGGL((k_chain_local_mixed<NRS, NBA, NBN, X2, KD>), grid, block, 0, s, a);
  if (int rc = hip_check(hipGetLastError(), "k_chain_local_mixed")) return rc;
  if (a.nseg > 1) {   // joins in the quad layout, two instances per workgroup
    const dim3 gj(chain_grid_x(a), a.nseg - 1, (B + 1) / 2);
    hipLaunchKernelGGL((k_chain_join<NRS, 2>), gj, block, 0, s, a);
    if (int rc = hip_check(hipGetLastError(), "k_chain_join")) return rc;
  }
  return SDDC_OK;
}

template <int NRS>
static int dispatch_nb(const ChainArgs& a, int NB, int B, hipStream_t s) {
  switch (NB) {
    case 1: return launch_chain_t<NRS, 1>(a, B, s);
    case 2: return launch_chain_t<NRS, 2>(a, B, s);
    case 3: return launch_chain_mixed_t<NRS, 2, 4, false, 2>(a, B, s);
    case 5: return launch_chain_mixed_t<NRS, 2, 4, true, 3>(a, B, s);
    case 6: return launch_chain_mixed_t<NRS, 3, 4, true, 2>(a, B, s);
    case 7: return launch_chain_mixed_t<NRS, 2, 4, true, 4>(a, B, s);
    case 8: return launch_chain_mixed_t<NRS, 2, 4, true, 6>(a, B, s);
    case 9: return launch_chain_mixed_t<NRS, 1, 4, true, 6>(a, B, s);
    case 10: return launch_chain_mixed_t<NRS, 1, 4, false, 4>(a, B, s);
    case 4: return launch_chain_t<NRS, 4>(a, B, s);
    default: return fail(SDDC_ERR_INVALID, "unsupported instances per workgroup");
  }
}

template <int NRS>
static int dispatch_mfma_nrs(const ChainArgs& a, int B, hipStream_t s) {
  return launch_chain_mfma_t<NRS>(a, B, s);
}

static int dispatch_chain(const ChainArgs& a, int nrs, int NB, int B, hipStream_t s) {
  if (NB == 16) {
    switch (nrs) {
      case 8: return dispatch_mfma_nrs<8>(a, B, s);
      case 16: return dispatch_mfma_nrs<16>(a, B, s);
      case 24: return dispatch_mfma_nrs<24>(a, B, s);
      case 32: return dispatch_mfma_nrs<32>(a, B, s);
      case 40: return dispatch_mfma_nrs<40>(a, B, s);
      case 48: return dispatch_mfma_nrs<48>(a, B, s);
      case 56: return dispatch_mfma_nrs<56>(a, B, s);
      case 64: return dispatch_mfma_nrs<64>(a, B, s);
      default: return fail(SDDC_ERR_UNSUPPORTED, "unsupported radial size");
    }
  }
  switch (nrs) {
    case 8: return dispatch_nb<8>(a, NB, B, s);
    case 16: return dispatch_nb<16>(a, NB, B, s);
    case 24: return dispatch_nb<24>(a, NB, B, s);
    case 32: return dispatch_nb<32>(a, NB, B, s);
    case 40: return dispatch_nb<40>(a, NB, B, s);
    case 48: return dispatch_nb<48>(a, NB, B, s);
    case 56: return dispatch_nb<56>(a, NB, B, s);
    case 64: return dispatch_nb<64>(a, NB, B, s);
    default: return fail(SDDC_ERR_UNSUPPORTED, "unsupported radial size");
  }
}

int launch_solve_ex(const Plan& p, int B, const double* G, long long gstride,
                    const long long goff[3], double* X, long long xstride,
                    const long long xoff[3], double* Pout, int fields_mask, hipStream_t s) {
  if (B <= 0) return SDDC_OK;
  if (B > p.ex_batch) return fail(SDDC_ERR_INVALID, "batch exceeds the plan's solve workspace");
  ChainArgs a;
  a.G = G;
  a.X = X;
  a.P = Pout;
  for (int f = 0; f < 3; ++f) {
    a.Lm[f] = p.chain_Lm[f];
    a.Lq[f] = p.chain_L[f];
    a.Qq[f] = p.chain_Q[f];
    a.Eq[f] = p.chain_E[f];
    a.goff[f] = goff[f];
    a.xoff[f] = xoff[f];
  }
  a.DLq = p.chain_DL;
  a.D2q = p.chain_D2;
  a.DLm = p.chain_DLm;
  a.vecs = p.vecs;
  a.EX = p.chain_EX;
  a.scratch = p.solve_scratch;
  a.gstride = gstride;
  a.xstride = xstride;
  a.K = p.K;
  a.nr = p.nr;
  a.B = B;
  a.m = p.seg_m;
  a.nseg = p.nseg;
  a.symmetric = p.symmetric;
  a.fields_mask = fields_mask;
  a.dtf[0] = p.Pr * p.dt;
  a.dtf[1] = p.dt;
  a.dtf[2] = p.Tau * p.dt;   // NAB2 of S uses Tau*dt (Main.py:281)
  if (p.nab2_modal && (fields_mask & 6)) {
    ModalArgs m;
    m.G = G;
    m.X = X;
    m.gstride = gstride;
    m.xstride = xstride;
    for (int f = 0; f < 3; ++f) {
      m.goff[f] = goff[f];
      m.xoff[f] = xoff[f];
      m.h[f] = a.dtf[f];
    }
    m.K = p.K;
    m.nr = p.nr;
    m.symmetric = p.symmetric;
    m.field0 = (fields_mask & 2) ? 1 : 2;
    const int nf = ((fields_mask >> 1) & 1) + ((fields_mask >> 2) & 1);
    if (int rc = dispatch_modal(m, nf, B, p, s)) return rc;
    a.fields_mask = fields_mask & 1;
    if (!a.fields_mask) return SDDC_OK;
  }
  // untuned plans: widths of one bit group (1-4 and 10 do the same
  // arithmetic), so the results do not depend on the batch size of a call.
  // From 64 instances width 3 (two A4 and four T/S instances per
  // workgroup): width 4 (four of each) needs one wavefront per SIMD from
  // N_r = 40 on and was 11-13% slower in the tuning runs at 64 and 256.
  int NB = B >= 64 ? 3 : (B >= 8 ? 2 : 1);
  if (p.tune_solve_nbi > 0) NB = p.tune_solve_nbi;   // 16: MFMA chain kernel
  return dispatch_chain(a, p.NRS, NB, B, s);
}

int launch_solve(const Plan& p, int B, const double* G, double* Xout, double* Pout,
                 int fields_mask, int /*mask_out*/, hipStream_t s) {
  const long long N = (long long)p.K * p.nr;
  const long long off[3] = {0, N, 2 * N};
  return launch_solve_ex(p, B, G, 3 * N, off, Xout, 3 * N, off, Pout, fields_mask, s);
}

// ---------------------------------------------------------------------------
// Host side: packing and the segment maps (built once per plan, float64).
// ---------------------------------------------------------------------------
// Quad packing of an nrows x ncols row-major matrix (leading dimension lda)
// into an NRP x NC tile (zero padded).
static void pack_quad(const double* A, int nrows, int ncols, int lda, int NRP, int NC,
                      double* out) {
  const int NQ = NC / 8;
  for (int w = 0; w < NRP / 16; ++w)
    for (int q = 0; q < NQ; ++q)
      for (int l = 0; l < 64; ++l)
        for (int h = 0; h < 2; ++h) {
          const int r = 16 * w + l / 4, c = (l & 3) + 8 * q + 4 * h;
          out[((size_t)(w * NQ + q) * 64 + l) * 2 + h] =
              (r < nrows && c < ncols) ? A[(size_t)r * lda + c] : 0.0;
        }
}

// C (m x n) = A (m x k) B (k x n), row-major.
static void matmul(const double* A, const double* B, double* C, int m, int k, int n) {
  for (int i = 0; i < m; ++i) {
    double* c = C + (size_t)i * n;
    for (int j = 0; j < n; ++j) c[j] = 0.0;
    for (int q = 0; q < k; ++q) {
      const double av = A[(size_t)i * k + q];
      const double* b = B + (size_t)q * n;
      for (int j = 0; j < n; ++j) c[j] += av * b[j];
    }
  }
}

// An nr x 2nr state matrix [F | B] packed as NRP x 2NRS, with columns of the
// second half starting at NRS (the kernels' state layout).
static void pack_state(const double* A, int nr, int NRS, int NRP, double* out) {
  std::vector<double> wide((size_t)nr * 2 * NRS, 0.0);
  for (int i = 0; i < nr; ++i)
    for (int c = 0; c < nr; ++c) {
      wide[(size_t)i * 2 * NRS + c] = A[(size_t)i * 2 * nr + c];
      wide[(size_t)i * 2 * NRS + NRS + c] = A[(size_t)i * 2 * nr + nr + c];
    }
  pack_quad(wide.data(), nr, 2 * NRS, 2 * NRS, NRP, 2 * NRS, out);
}

// MFMA packing of an nrows x ncols row-major matrix into an NRP x NC tile:
// element (16w + l%16, 4ks + l/16) at ((w*(NC/8) + ks/2)*64 + l)*2 + ks%2, so
// that lane l of wavefront w loads its A fragments of k steps 2kp, 2kp+1 with
// one 16-byte load (coalesced across the wavefront).
static void pack_mfma(const double* A, int nrows, int ncols, int lda, int NRP, int NC,
                      double* out) {
  const int NKP = NC / 8;
  for (int w = 0; w < NRP / 16; ++w)
    for (int kp = 0; kp < NKP; ++kp)
      for (int l = 0; l < 64; ++l)
        for (int h = 0; h < 2; ++h) {
          const int r = 16 * w + (l & 15), c = 4 * (2 * kp + h) + (l >> 4);
          out[((size_t)(w * NKP + kp) * 64 + l) * 2 + h] =
              (r < nrows && c < ncols) ? A[(size_t)r * lda + c] : 0.0;
        }
}

// Rows per segment for chains of n rows.  Small batches are latency bound
// (depth m + n/m, smallest near m = sqrt(n)); mid-size batches trade depth for
// fewer corrections (m = 2 sqrt(n)); large batches fill the device without
// segments, which avoid the corrections' extra work (measured at n = 256 on
// MI355X).  SDDC_CHAIN_SEGMENT overrides.  Fixed per plan, so that the bits
// of a result do not depend on the batch it was computed in.
static int default_segment(int n, int max_batch) {
  if (const char* env = std::getenv("SDDC_CHAIN_SEGMENT")) {
    const int m = std::atoi(env);
    if (m > 0) return m;
  }
  if (max_batch > 96) return n;
  const double c = max_batch <= 4 ? 1.0 : 2.0;
  int m = (int)std::ceil(c * std::sqrt((double)n));
  m = ((m + kDepth - 1) / kDepth) * kDepth;
  return m < 8 ? 8 : m;
}

template <typename T>
static int upload_vec(T** dst, const std::vector<T>& v) {
  SDDC_HIP(hipMalloc(dst, v.size() * sizeof(T)));
  SDDC_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return SDDC_OK;
}

int build_chain(Plan& p, const sddc_operators* ops, int max_batch, int segment) {
  const int K = p.K, nr = p.nr, NRS = p.NRS, NRP = p.NRP;
  const size_t BLK = (size_t)NRP * NRS, NN = (size_t)nr * nr;
  const int n = K / 2;   // rows per chain (K is even)
  int m = segment > 0 ? segment : default_segment(n, max_batch);
  if (m > n) m = n;
  p.seg_m = m;
  p.nseg = (n + m - 1) / m;
  const double* inv[3] = {ops->Linv_A4, ops->Linv_T, ops->Linv_S};
  const int nfb = p.nab2_modal ? 1 : 3;   // fields solved with per-mode blocks
  std::vector<double> buf((size_t)K * BLK);
  for (int f = 0; f < nfb; ++f) {
    for (int jj = 0; jj < K; ++jj)
      pack_quad(inv[f] + jj * NN, nr, nr, nr, NRP, NRS, buf.data() + jj * BLK);
    if (int rc = upload_vec(&p.chain_L[f], buf)) return rc;
  }
  if (p.nab2_modal) {   // lam | V | V^-1, rows padded to NRS
    const double* md[3] = {nullptr, ops->modes_T, ops->modes_S};
    for (int f = 1; f < 3; ++f) {
      std::vector<double> m((size_t)NRS + (size_t)nr * NRS + (size_t)NRS * NRS, 0.0);
      for (int k = 0; k < NRS; ++k) m[k] = k < nr ? md[f][k] : 1.0;
      for (int r = 0; r < nr; ++r)
        for (int k = 0; k < nr; ++k) {
          m[NRS + (size_t)r * NRS + k] = md[f][nr + (size_t)r * nr + k];
          m[NRS + (size_t)nr * NRS + (size_t)r * NRS + k] = md[f][nr + NN + (size_t)r * nr + k];
        }
      if (int rc = upload_vec(&p.nab2_modes[f], m)) return rc;
    }
  }
  // D2 Linv_j per mode (float64 on the host): the A4 chains carry D2 f_e
  std::vector<double> dl((size_t)K * NN);
  for (int jj = 0; jj < K; ++jj) matmul(ops->D2_A4, inv[0] + jj * NN, dl.data() + jj * NN, nr, nr, nr);
  for (int jj = 0; jj < K; ++jj)
    pack_quad(dl.data() + jj * NN, nr, nr, nr, NRP, NRS, buf.data() + jj * BLK);
  if (int rc = upload_vec(&p.chain_DL, buf)) return rc;
  {   // D2 itself, for the two-exchange A4 chains
    std::vector<double> d2b((size_t)BLK);
    pack_quad(ops->D2_A4, nr, nr, nr, NRP, NRS, d2b.data());
    if (int rc = upload_vec(&p.chain_D2, d2b)) return rc;
  }
  if (p.tune_solve_nbi == 16) {   // operands of the MFMA chain kernel
    for (int f = 0; f < nfb; ++f) {
      for (int jj = 0; jj < K; ++jj)
        pack_mfma(inv[f] + jj * NN, nr, nr, nr, NRP, NRS, buf.data() + jj * BLK);
      if (int rc = upload_vec(&p.chain_Lm[f], buf)) return rc;
    }
    for (int jj = 0; jj < K; ++jj)
      pack_mfma(dl.data() + jj * NN, nr, nr, nr, NRP, NRS, buf.data() + jj * BLK);
    if (int rc = upload_vec(&p.chain_DLm, buf)) return rc;
  }
  if (p.nseg == 1) return SDDC_OK;   // one segment per chain: no joins

  // NAB2 (fields 1, 2).  State b; T maps the entry state to b at the row:
  // Q = -c_j Linv_j T, T <- T + 2 dtf j Q; E = T after the segment.
  const double dts[3] = {p.Pr * p.dt, p.dt, p.Tau * p.dt};
  std::vector<double> T(NN), Q(NN), LT(NN);
  for (int f = 1; f < nfb; ++f) {
    std::vector<double> qpk((size_t)K * BLK, 0.0), epk((size_t)2 * p.nseg * BLK, 0.0);
    for (int chain = 0; chain < 2; ++chain) {
      const int start = K - 1 - chain;
      for (int s = 0; s < p.nseg; ++s) {
        std::fill(T.begin(), T.end(), 0.0);
        for (int i = 0; i < nr; ++i) T[(size_t)i * nr + i] = 1.0;
        const int k0 = s * m, k1 = std::min(k0 + m, n);
        for (int k = k0; k < k1; ++k) {
          const int j = start - 2 * k, jj = K - 1 - j;
          const double cj = (j == 0) ? 0.5 : 1.0;
          matmul(inv[f] + jj * NN, T.data(), LT.data(), nr, nr, nr);
          for (size_t e = 0; e < NN; ++e) Q[e] = -cj * LT[e];
          pack_quad(Q.data(), nr, nr, nr, NRP, NRS, qpk.data() + jj * BLK);
          const double alpha = 2.0 * dts[f] * j;
          for (size_t e = 0; e < NN; ++e) T[e] += alpha * Q[e];
        }
        pack_quad(T.data(), nr, nr, nr, NRP, NRS, epk.data() + ((size_t)chain * p.nseg + s) * BLK);
      }
    }
    if (int rc = upload_vec(&p.chain_Q[f], qpk)) return rc;
    if (int rc = upload_vec(&p.chain_E[f], epk)) return rc;
  }

  // A4 (field 0).  State (f_e, bf_e), 2nr; Tf, Tb (nr x 2nr) map the entry
  // state to f_e and bf_e at the row.  With j = row + 1, bj = -j(j+1),
  // bjt = -2j:  M T = dt bjt (D2 Tf + bj IR4 Tf + IR4 Tb) - bjt IR2 Tf,
  // Q = Linv M T,  Tb <- Tb + bj Q + bjt Tf,  Tf <- Tf + Q.
  {
    const size_t NW = (size_t)nr * 2 * nr, BLK2 = 2 * BLK;
    const double dta = dts[0];
    std::vector<double> Tf(NW), Tb(NW), D2T(NW), MT(NW), Q2(NW);
    std::vector<double> qpk((size_t)K * BLK2, 0.0), epk((size_t)2 * p.nseg * 2 * BLK2, 0.0);
    for (int chain = 0; chain < 2; ++chain) {
      const int start = K - 1 - chain;
      for (int s = 0; s < p.nseg; ++s) {
        std::fill(Tf.begin(), Tf.end(), 0.0);
        std::fill(Tb.begin(), Tb.end(), 0.0);
        for (int i = 0; i < nr; ++i) {
          Tf[(size_t)i * 2 * nr + i] = 1.0;
          Tb[(size_t)i * 2 * nr + nr + i] = 1.0;
        }
        const int k0 = s * m, k1 = std::min(k0 + m, n);
        for (int k = k0; k < k1; ++k) {
          const int row = start - 2 * k, jj = K - 1 - row;
          const double jp = row + 1.0, bj = -jp * (jp + 1.0), bjt = -2.0 * jp;
          matmul(ops->D2_A4, Tf.data(), D2T.data(), nr, nr, 2 * nr);
          for (int i = 0; i < nr; ++i) {
            const double ir4 = ops->IR4_A4[i], ir2 = ops->IR2_A4[i];
            for (int c = 0; c < 2 * nr; ++c) {
              const size_t e = (size_t)i * 2 * nr + c;
              MT[e] = dta * bjt * (D2T[e] + bj * ir4 * Tf[e] + ir4 * Tb[e]) - bjt * ir2 * Tf[e];
            }
          }
          matmul(inv[0] + jj * NN, MT.data(), Q2.data(), nr, nr, 2 * nr);
          pack_state(Q2.data(), nr, NRS, NRP, qpk.data() + jj * BLK2);
          for (size_t e = 0; e < NW; ++e) {
            Tb[e] += bj * Q2[e] + bjt * Tf[e];
            Tf[e] += Q2[e];
          }
        }
        double* eo = epk.data() + ((size_t)chain * p.nseg + s) * 2 * BLK2;
        pack_state(Tf.data(), nr, NRS, NRP, eo);
        pack_state(Tb.data(), nr, NRS, NRP, eo + BLK2);
      }
    }
    if (int rc = upload_vec(&p.chain_Q[0], qpk)) return rc;
    if (int rc = upload_vec(&p.chain_E[0], epk)) return rc;
  }
  return SDDC_OK;
}

void free_chain(Plan& p) {
  for (int f = 0; f < 3; ++f) {
    if (p.chain_L[f]) (void)hipFree(p.chain_L[f]);
    if (p.chain_Q[f]) (void)hipFree(p.chain_Q[f]);
    if (p.chain_E[f]) (void)hipFree(p.chain_E[f]);
    p.chain_L[f] = p.chain_Q[f] = p.chain_E[f] = nullptr;
  }
  if (p.chain_DL) (void)hipFree(p.chain_DL);
  p.chain_DL = nullptr;
  if (p.chain_D2) (void)hipFree(p.chain_D2);
  p.chain_D2 = nullptr;
  for (int f = 0; f < 3; ++f) {
    if (p.chain_Lm[f]) (void)hipFree(p.chain_Lm[f]);
    p.chain_Lm[f] = nullptr;
  }
  if (p.chain_DLm) (void)hipFree(p.chain_DLm);
  p.chain_DLm = nullptr;
  for (int f = 0; f < 3; ++f) {
    if (p.nab2_modes[f]) (void)hipFree(p.nab2_modes[f]);
    p.nab2_modes[f] = nullptr;
  }
}

}  // namespace sddc
