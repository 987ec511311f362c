// Internal definitions shared by the HIP translation units of libsddc.
//
// Data layout in HBM (one instance; a batch of B instances of one size class
// is B such blocks back to back):
//   X      [3][K][nr]   float64  state psi|T|S, mode-major (reference layout,
//                                Matrix_Operators.py:529-627)
//   P      [K][nr]      float64  same-parity suffix sums of psi rows
//   L      [3][K][nr]   float64  linear part of the implicit right-hand side
//   G      [3][K][nr]   float64  full right-hand side of the three solves
//   slabs  [7][nr][K]   real_t   theta rows feeding the inverse transforms
//   F      [4][nr][K]   real_t   projected products (sine Q, sine R, cos T, cos S)
// with K = N_fm modes, nr = N_r - 1 interior Chebyshev points, and real_t the
// nonlinear-pipeline precision (float64 plain, float32 mixed).
#pragma once
#include <atomic>
#include <mutex>

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "sddc.h"

namespace sddc {

constexpr int kMaxRadices = 24;
constexpr int kTile = 16;          // modes per workgroup in the spectral stages
constexpr int kMaxNr = 64;         // largest supported interior point count

// Waits for every memory operation in flight.  Issued once after a loop's
// prologue (its first fetches and the loop-invariant operands): without it
// the compiler's wait-count analysis merges the prologue's issue order into
// the loop header and waits, on every trip, for loads issued for a later
// trip, which serialises the loop on memory latency.
__device__ __forceinline__ void drain_loads() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);   // vmcnt 0, expcnt 0, lgkmcnt 0
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Workgroup barrier for data exchanged through LDS: the wavefront's own LDS
// accesses are complete (lgkmcnt 0) before s_barrier.  No vmcnt wait, so
// global loads in flight stay in flight.  Signal fences keep the compiler
// from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xC07F);   // vmcnt 63 (none), expcnt 7 (none), lgkmcnt 0
  __builtin_amdgcn_s_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

template <typename R>
struct cplx {
  R x, y;
};

template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cplx<R> csub(cplx<R> a, cplx<R> b) { return {a.x - b.x, a.y - b.y}; }

// Radix list of an FFT length (passed by value as a kernel argument).
struct FftShape {
  int L;
  int npass;
  int radix[kMaxRadices];
};

// Device tables of one transform length L: forward twiddles exp(-2 pi i t/L)
// and the half-sample rotation exp(i pi k/(2L)), both rounded from float64.
template <typename R>
struct FftTables {
  const cplx<R>* tw;
  const cplx<R>* rot;
};

struct TableSet {
  int L = 0;
  FftShape shape{};
  double2* tw64 = nullptr;
  double2* rot64 = nullptr;
  float2* tw32 = nullptr;
  float2* rot32 = nullptr;
  // the same twiddles laid out per pass of the shape (see ip_twiddle_offset):
  // a pass reads its entries at consecutive addresses
  double2* twp64 = nullptr;
  float2* twp32 = nullptr;
};

// Per-size-class operator set and workspace.
struct Plan {
  int K = 0, N_r = 0, nr = 0, NRP = 0, M = 0;
  double d = 0, dt = 0, Pr = 0, Tau = 0;
  int precision = 0;  // 0: float64 throughout; 1: float32 nonlinear pipeline
  int symmetric = 0;
  int device = 0;
  int max_batch = 0;
  // operators (device)
  double* Dfull = nullptr;     // (N_r+1)^2 row-major
  double* R = nullptr;         // N_r+1
  double* mfma64 = nullptr;    // MFMA-tiled Dr | D2r2 | D2s   (3 * NRP*NRP)
  float* mfma32 = nullptr;     // MFMA-tiled Dr | D2r2          (2 * NRP*NRP)
  double* vecs = nullptr;      // 8 vectors of length NRP (see VecId)
  // implicit solves: quad-packed operator data (layout in sddc_solve.hip)
  int NRS = 0;                 // nr rounded up to a multiple of 8
  int seg_m = 0, nseg = 0;     // rows per chain segment, segments per chain
  double* chain_L[3] = {nullptr, nullptr, nullptr};   // block inverses, K x NRP x NRS
  double* chain_DL = nullptr;                         // D2 Linv_j of the A4 solve, per mode
  double* chain_D2 = nullptr;                         // D2 of the A4 solve (one block)
  double* chain_Q[3] = {nullptr, nullptr, nullptr};   // row corrections (nseg > 1)
  double* chain_E[3] = {nullptr, nullptr, nullptr};   // segment exit maps (nseg > 1)
  double* chain_Lm[3] = {nullptr, nullptr, nullptr};  // MFMA-packed block inverses
  double* chain_DLm = nullptr;                        // MFMA-packed D2 Linv_j (A4)
  // NAB2 solves in the eigenbasis of r^2 - h r^2 Lap (fields 1, 2), when set:
  // per field lam [NRS] (padding 1), V [nr][NRS], V^-1 [NRS][NRS] (padding 0)
  int nab2_modal = 0;
  double* nab2_modes[3] = {nullptr, nullptr, nullptr};
  TableSet tabM, tabKE;
  // workspace (device), sized for max_batch instances
  double* P = nullptr;
  double* L = nullptr;
  double* G = nullptr;
  double* Xa = nullptr;
  double* Xb = nullptr;
  double* Pa = nullptr;
  void* slabs = nullptr;
  void* slabs2 = nullptr;
  void* F = nullptr;
  double* rowval = nullptr;
  double* params = nullptr;    // 2*max_batch: Ra, Ra_s staging
  double* solve_scratch = nullptr;   // 4N + 256 doubles, sink of padded solve stores
  double* chain_EX = nullptr;        // zero-entry segment exit states
  int ex_batch = 0;                  // instances chain_EX holds
  int tune_solve_nbi = 0;            // instances per solve workgroup, 0 = automatic
  int tune_sub_batches = 1;          // parts of a batch on forked streams in graphs
  int tune_spectral_nb = 0;          // instances per spectral workgroup, 0 = automatic
  size_t ws_bytes = 0;
};

// Radial vectors stored in Plan::vecs, each NRP long (zero padded):
// 1/r^2, 1/r^4, A_T/r^2, r^2, R_1^2/r^2, diag IR2 and IR4 of the A4 solve, 1/r.
enum VecId { V_IR2 = 0, V_IR4, V_DT0, V_RSQ, V_GR, V_IR2A4, V_IR4A4, V_IR, V_COUNT };

// error plumbing
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_check(hipError_t e, const char* what);
#define SDDC_HIP(call)                                  \
  do {                                                  \
    int _rc = ::sddc::hip_check((call), #call);         \
    if (_rc != SDDC_OK) return _rc;                     \
  } while (0)

// The device-state lock.  Held for the whole of a HIP graph capture and by
// every library call that allocates, frees, copies synchronously to or from
// the device or synchronises the device (plan create/reserve/destroy,
// Krylov create/destroy, graph destroy, transform tables, kernel attribute
// changes).  Such calls from another thread during a capture can invalidate
// it (hipErrorStreamCaptureInvalidated at hipStreamEndCapture) even in
// thread-local capture mode; under the lock they wait for the capture to end.
std::recursive_mutex& device_state_mutex();

int build_tables(TableSet& t, int L);
void free_tables(TableSet& t);
FftShape make_shape(int L);

// launchers implemented in the kernel translation units
int launch_suffix(const Plan& p, int B, const double* X, double* P, hipStream_t s);
int launch_spectral(const Plan& p, int B, const double* X, const double* P,
                    const double* Ra, const double* Ras, int want_linear, double* L,
                    void* slabs, int slab_precision, double* theta_out, double* slab9,
                    hipStream_t s);
int launch_nonlinear(const Plan& p, int B, const void* slabs, const void* slabs2, void* F,
                     hipStream_t s);
int launch_rhs(const Plan& p, int B, const void* F, const double* L, int have_nonlin,
               double* G, hipStream_t s);
int launch_solve(const Plan& p, int B, const double* G, double* Xout, double* Pout,
                 int fields_mask, int mask_out, hipStream_t s);
int build_chain(Plan& p, const sddc_operators* ops, int max_batch, int segment);
void free_chain(Plan& p);
int launch_mask(const Plan& p, int B, const double* X, double* out, hipStream_t s);
int launch_transform(const TableSet& t, int kind, int rows, int len_in, int n_out,
                     const double* in, double* out, hipStream_t s);
int launch_ke_rows(const Plan& p, int B, const void* slabs64, double* rowval, hipStream_t s);
int launch_diag_final(const Plan& p, int B, const double* X, const double* rowval,
                      double* out, hipStream_t s);
int nonlinear_lds_bytes(const Plan& p, int precision, int dfx);
// Raise a kernel's dynamic-LDS limit to `bytes` (once per size increase).
int ensure_dynamic_lds(const void* fn, size_t bytes, std::atomic<size_t>& current);

}  // namespace sddc

// The opaque handle of include/sddc.h.
struct sddc_plan {
  sddc::Plan p;
  int live_graphs = 0;   // graphs holding this plan's workspace pointers (under the lock)
};

