// C ABI of libsddc: plan lifecycle, operator upload and the entry points
// declared in include/sddc.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>

#include <vector>

#include "sddc_fft.h"
#include "sddc_internal.h"


struct sddc_graph {
  std::vector<sddc_plan*> plans;   // plans whose workspace the graph holds (one count each)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t own = nullptr;
  std::vector<hipStream_t> side;   // the capture's lane streams (at most one per hardware queue)
  std::vector<hipEvent_t> events;
};

namespace sddc {

static thread_local std::string g_last_error;

std::recursive_mutex& device_state_mutex() {
  static std::recursive_mutex mu;
  return mu;
}

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}
int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return SDDC_OK;
  char buf[512];
  snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
  set_error(buf);
  return SDDC_ERR_HIP;
}

FftShape make_shape(int L) {
  FftShape s{};
  s.L = L;
  s.npass = 0;
  int n = L;
  const int pref[4] = {4, 2, 3, 5};
  for (int r : pref)
    while (n % r == 0 && s.npass < kMaxRadices) {
      s.radix[s.npass++] = r;
      n /= r;
    }
  for (int p = 7; n > 1 && s.npass < kMaxRadices; p += 2)
    while (n % p == 0) {
      s.radix[s.npass++] = p;
      n /= p;
    }
  if (n != 1) s.npass = -1;
  return s;
}

int build_tables(TableSet& t, int L) {
  t.L = L;
  t.shape = make_shape(L);
  if (t.shape.npass < 0) return fail(SDDC_ERR_UNSUPPORTED, "transform length has too many factors");
  std::vector<double2> tw(L), rot(L);
  std::vector<float2> tw32(L), rot32(L);
  const long double PI = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < L; ++k) {
    // exp(-2 pi i k/L) and exp(i pi k/(2L)), evaluated in extended precision
    const long double a = -2.0L * PI * (long double)k / (long double)L;
    const long double b = PI * (long double)k / (2.0L * (long double)L);
    tw[k] = make_double2((double)cosl(a), (double)sinl(a));
    rot[k] = make_double2((double)cosl(b), (double)sinl(b));
    tw32[k] = make_float2((float)tw[k].x, (float)tw[k].y);
    rot32[k] = make_float2((float)rot[k].x, (float)rot[k].y);
  }
  // per-pass layout of the same values for the in-place transforms (entry
  // L-1 is unused); its schedule is make_shape's for lengths of the factors
  // 2, 3 and 5, the only ones the in-place kernels take
  std::vector<double2> twp(L, make_double2(1.0, 0.0));
  std::vector<float2> twp32(L, make_float2(1.0f, 0.0f));
  if (ip_npass(L) > 0 && ip_twiddle_offset(L, ip_npass(L)) == L - 1) {
    int ns = 1;
    for (int q = 0; q < ip_npass(L); ++q) {
      const int r = ip_radix(L, q), st = L / (ns * r), off = ip_twiddle_offset(L, q);
      for (int s = 1; s < r; ++s)
        for (int k = 0; k < ns; ++k) {
          twp[off + (s - 1) * ns + k] = tw[s * k * st];
          twp32[off + (s - 1) * ns + k] = tw32[s * k * st];
        }
      ns *= r;
    }
  }
  SDDC_HIP(hipMalloc(&t.tw64, L * sizeof(double2)));
  SDDC_HIP(hipMalloc(&t.rot64, L * sizeof(double2)));
  SDDC_HIP(hipMalloc(&t.tw32, L * sizeof(float2)));
  SDDC_HIP(hipMalloc(&t.rot32, L * sizeof(float2)));
  SDDC_HIP(hipMalloc(&t.twp64, L * sizeof(double2)));
  SDDC_HIP(hipMalloc(&t.twp32, L * sizeof(float2)));
  SDDC_HIP(hipMemcpy(t.tw64, tw.data(), L * sizeof(double2), hipMemcpyHostToDevice));
  SDDC_HIP(hipMemcpy(t.rot64, rot.data(), L * sizeof(double2), hipMemcpyHostToDevice));
  SDDC_HIP(hipMemcpy(t.tw32, tw32.data(), L * sizeof(float2), hipMemcpyHostToDevice));
  SDDC_HIP(hipMemcpy(t.rot32, rot32.data(), L * sizeof(float2), hipMemcpyHostToDevice));
  SDDC_HIP(hipMemcpy(t.twp64, twp.data(), L * sizeof(double2), hipMemcpyHostToDevice));
  SDDC_HIP(hipMemcpy(t.twp32, twp32.data(), L * sizeof(float2), hipMemcpyHostToDevice));
  return SDDC_OK;
}

void free_tables(TableSet& t) {
  if (t.tw64) (void)hipFree(t.tw64);
  if (t.rot64) (void)hipFree(t.rot64);
  if (t.tw32) (void)hipFree(t.tw32);
  if (t.rot32) (void)hipFree(t.rot32);
  if (t.twp64) (void)hipFree(t.twp64);
  if (t.twp32) (void)hipFree(t.twp32);
  t = TableSet{};
}

int launch_solve_ex(const Plan& p, int B, const double* G, long long gstride,
                    const long long goff[3], double* X, long long xstride,
                    const long long xoff[3], double* Pout, int fields_mask, hipStream_t s);
int launch_rhs_scaled(const Plan& p, int B, const void* F, const double* L, double scale,
                      double* G, hipStream_t s);

// MFMA operand tiling: T[it][kk][lane] = A[16*it + lane%16][4*kk + lane/16].
template <typename T>
static void tile_mfma(const double* A, int nr, int NRP, T* out) {
  const int NT = NRP / 16, NK = NRP / 4;
  for (int it = 0; it < NT; ++it)
    for (int kk = 0; kk < NK; ++kk)
      for (int lane = 0; lane < 64; ++lane) {
        const int r = 16 * it + (lane & 15), c = 4 * kk + (lane >> 4);
        const double v = (r < nr && c < nr) ? A[(size_t)r * nr + c] : 0.0;
        out[((size_t)it * NK + kk) * 64 + lane] = (T)v;
      }
}

template <typename T>
static int upload(T** dst, const T* src, size_t n) {
  SDDC_HIP(hipMalloc(dst, n * sizeof(T)));
  SDDC_HIP(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return SDDC_OK;
}

static void free_ws(Plan& p) {
  void* bufs[] = {p.P, p.L, p.G, p.Xa, p.Xb, p.Pa, p.slabs, p.slabs2, p.F, p.rowval, p.params,
                  p.solve_scratch, p.chain_EX};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  p.P = p.L = p.G = p.Xa = p.Xb = p.Pa = nullptr;
  p.slabs = p.slabs2 = p.F = nullptr;
  p.rowval = p.params = p.solve_scratch = p.chain_EX = nullptr;
  p.ex_batch = 0;
  p.ws_bytes = 0;
  p.max_batch = 0;
}

static int alloc_ws(Plan& p, int B) {
  if (B <= p.max_batch) return SDDC_OK;
  free_ws(p);
  const size_t N = (size_t)p.K * p.nr;
  const size_t rs = p.precision == SDDC_MIXED ? sizeof(float) : sizeof(double);
  const size_t slab_bytes = (size_t)B * 7 * N * (rs > sizeof(double) ? rs : sizeof(double));
  size_t total = 0;
  auto A = [&](void** ptr, size_t bytes) -> int {
    total += bytes;
    return hip_check(hipMalloc(ptr, bytes > 0 ? bytes : 8), "workspace");
  };
  int rc;
  if ((rc = A((void**)&p.P, B * N * 8))) return rc;
  if ((rc = A((void**)&p.Pa, B * N * 8))) return rc;
  if ((rc = A((void**)&p.L, B * 3 * N * 8))) return rc;
  if ((rc = A((void**)&p.G, B * 3 * N * 8))) return rc;
  if ((rc = A((void**)&p.Xa, B * 3 * N * 8))) return rc;
  if ((rc = A((void**)&p.Xb, B * 3 * N * 8))) return rc;
  // slabs hold real_t; slabs2 doubles as the float64 slab set of diagnostics
  if ((rc = A(&p.slabs, slab_bytes))) return rc;
  if ((rc = A(&p.slabs2, (size_t)B * 7 * N * 8))) return rc;
  if ((rc = A(&p.F, (size_t)B * 4 * N * rs))) return rc;
  if ((rc = A((void**)&p.rowval, (size_t)B * p.nr * 8))) return rc;
  if ((rc = A((void**)&p.params, (size_t)2 * B * 8))) return rc;
  // sink of padded instances and rows: X at [0, N), P at [N, 2N), lanes from 2N
  if ((rc = A((void**)&p.solve_scratch, (4 * N + 256) * 8))) return rc;
  // segment exit states of the solve, for B rounded up to whole workgroups
  const int exb = ((B + 3) / 4) * 4;
  if ((rc = A((void**)&p.chain_EX, (size_t)exb * 6 * p.nseg * 2 * p.NRP * 8))) return rc;
  p.ex_batch = exb;
  p.ws_bytes = total;
  p.max_batch = B;
  return SDDC_OK;
}

static int check_batch(Plan& p, int B) {
  if (B < 0) return fail(SDDC_ERR_INVALID, "negative batch");
  if (B > p.max_batch) return fail(SDDC_ERR_INVALID, "batch exceeds the plan's reserved size");
  return SDDC_OK;
}

}  // namespace sddc

using namespace sddc;

extern "C" {

static int static_scratch(int dev, unsigned char** out);

const char* sddc_last_error(void) { return g_last_error.c_str(); }
const char* sddc_version(void) { return "sddc 0.1 gfx950"; }

int sddc_plan_create(const sddc_operators* ops, int precision, int symmetric, int max_batch,
                     sddc_plan** out) {
  return sddc_plan_create_tuned(ops, precision, symmetric, max_batch, nullptr, out);
}

int sddc_plan_tuning(const sddc_plan* plan, sddc_tuning* out) {
  if (!plan || !out) return fail(SDDC_ERR_INVALID, "null argument");
  *out = sddc_tuning{};
  out->chain_segment = plan->p.seg_m;
  out->solve_instances = plan->p.tune_solve_nbi;
  out->sub_batches = plan->p.tune_sub_batches;
  out->nab2_solve = plan->p.nab2_modal ? 2 : 1;
  out->spectral_instances = plan->p.tune_spectral_nb;
  return SDDC_OK;
}

int sddc_plan_create_tuned(const sddc_operators* ops, int precision, int symmetric, int max_batch,
                           const sddc_tuning* tuning, sddc_plan** out) {
  if (!ops || !out) return fail(SDDC_ERR_INVALID, "null argument");
  *out = nullptr;
  std::lock_guard<std::recursive_mutex> lk(device_state_mutex());
  if (ops->N_fm % 2 != 0) {
    char buf[96];
    snprintf(buf, sizeof(buf), "The number of Fourier modes is not even %d", ops->N_fm);
    return fail(SDDC_ERR_ODD_NFM, buf);
  }
  if (ops->N_fm < 2 || ops->N_r < 3) return fail(SDDC_ERR_INVALID, "N_fm >= 2 and N_r >= 3 required");
  if (ops->N_r - 1 > kMaxNr) return fail(SDDC_ERR_UNSUPPORTED, "N_r - 1 above 64 is not supported");
  if (precision != SDDC_FP64 && precision != SDDC_MIXED)
    return fail(SDDC_ERR_INVALID, "precision must be SDDC_FP64 or SDDC_MIXED");
  const double* req[] = {ops->D, ops->R, ops->Dr, ops->D2s, ops->D2r2, ops->IR2v, ops->IR4v,
                         ops->DT0, ops->Rsq, ops->GR, ops->D2_A4, ops->IR2_A4, ops->IR4_A4,
                         ops->Linv_A4, ops->Linv_T, ops->Linv_S};
  for (const double* q : req)
    if (!q) return fail(SDDC_ERR_INVALID, "null operator pointer");
  sddc_plan* h = new sddc_plan();
  Plan& p = h->p;
  p.K = ops->N_fm;
  p.N_r = ops->N_r;
  p.nr = ops->N_r - 1;
  p.NRP = ((p.nr + 15) / 16) * 16;
  p.M = 3 * p.K / 2;
  p.d = ops->d;
  p.dt = ops->dt;
  p.Pr = ops->Pr;
  p.Tau = ops->Tau;
  p.precision = precision;
  p.symmetric = symmetric ? 1 : 0;
  int rc = SDDC_OK;
  auto bail = [&](int code) {
    sddc_plan_destroy(h);
    return code;
  };
  if (hipGetDevice(&p.device) != hipSuccess) return bail(fail(SDDC_ERR_HIP, "no HIP device"));
  // set-up of sddc_step_static: this device's address of the static scratch
  {
    unsigned char* base = nullptr;
    if ((rc = static_scratch(p.device, &base))) return bail(rc);
  }
  int lds_max = 0;
  (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, p.device);
  if ((size_t)nonlinear_lds_bytes(p, precision, 0) > (size_t)(lds_max > 0 ? lds_max : 65536))
    return bail(fail(SDDC_ERR_UNSUPPORTED, "N_fm too large for the LDS nonlinear kernel"));
  const int nr = p.nr, NRP = p.NRP, Np1 = p.N_r + 1;
  if ((rc = upload(&p.Dfull, ops->D, (size_t)Np1 * Np1))) return bail(rc);
  if ((rc = upload(&p.R, ops->R, (size_t)Np1))) return bail(rc);
  const size_t T = (size_t)NRP * NRP;
  std::vector<double> t64(3 * T);
  std::vector<float> t32(2 * T);
  tile_mfma<double>(ops->Dr, nr, NRP, t64.data());
  tile_mfma<double>(ops->D2r2, nr, NRP, t64.data() + T);
  tile_mfma<double>(ops->D2s, nr, NRP, t64.data() + 2 * T);
  tile_mfma<float>(ops->Dr, nr, NRP, t32.data());
  tile_mfma<float>(ops->D2r2, nr, NRP, t32.data() + T);
  if ((rc = upload(&p.mfma64, t64.data(), t64.size()))) return bail(rc);
  if ((rc = upload(&p.mfma32, t32.data(), t32.size()))) return bail(rc);
  std::vector<double> v((size_t)V_COUNT * NRP, 0.0);
  for (int i = 0; i < nr; ++i) {
    v[V_IR2 * NRP + i] = ops->IR2v[i];
    v[V_IR4 * NRP + i] = ops->IR4v[i];
    v[V_DT0 * NRP + i] = ops->DT0[i];
    v[V_RSQ * NRP + i] = ops->Rsq[i];
    v[V_GR * NRP + i] = ops->GR[i];
    v[V_IR2A4 * NRP + i] = ops->IR2_A4[i];
    v[V_IR4A4 * NRP + i] = ops->IR4_A4[i];
    v[V_IR * NRP + i] = 1. / ops->R[i + 1];
  }
  if ((rc = upload(&p.vecs, v.data(), v.size()))) return bail(rc);
  p.NRS = ((nr + 7) / 8) * 8;
  if (tuning && !((tuning->solve_instances >= 0 && tuning->solve_instances <= 10) ||
                  tuning->solve_instances == 16))
    return bail(fail(SDDC_ERR_INVALID, "solve_instances must be 0-10 or 16"));
  if (tuning && tuning->chain_segment < 0)
    return bail(fail(SDDC_ERR_INVALID, "chain_segment must be >= 0"));
  if (tuning && (tuning->sub_batches < 0 || tuning->sub_batches > 8))
    return bail(fail(SDDC_ERR_INVALID, "sub_batches must be in [0, 8]"));
  if (tuning && (tuning->nab2_solve < 0 || tuning->nab2_solve > 2))
    return bail(fail(SDDC_ERR_INVALID, "nab2_solve must be 0, 1 or 2"));
  if (tuning && (tuning->spectral_instances < 0 || tuning->spectral_instances > 4))
    return bail(fail(SDDC_ERR_INVALID, "spectral_instances must be in [0, 4]"));
  p.tune_spectral_nb = tuning ? tuning->spectral_instances : 0;
  {
    const bool have_modes = ops->modes_T && ops->modes_S;
    const int want = tuning ? tuning->nab2_solve : 0;
    if (want == 2 && !have_modes)
      return bail(fail(SDDC_ERR_INVALID, "nab2_solve 2 needs the operator set's modes_T and modes_S"));
    p.nab2_modal = want == 2 ? 1 : 0;
  }
  p.tune_solve_nbi = tuning ? tuning->solve_instances : 0;
  p.tune_sub_batches = (tuning && tuning->sub_batches > 0) ? tuning->sub_batches : 1;
  if ((rc = build_chain(p, ops, max_batch > 0 ? max_batch : 1,
                        tuning ? tuning->chain_segment : 0)))
    return bail(rc);
  if ((rc = build_tables(p.tabM, p.M))) return bail(rc);
  if ((rc = build_tables(p.tabKE, 3 * p.K))) return bail(rc);
  if ((rc = alloc_ws(p, max_batch > 0 ? max_batch : 1))) return bail(rc);
  *out = h;
  return SDDC_OK;
}

int sddc_plan_reserve(sddc_plan* plan, int max_batch) {
  if (!plan) return fail(SDDC_ERR_INVALID, "null plan");
  std::lock_guard<std::recursive_mutex> lk(device_state_mutex());
  if (max_batch <= plan->p.max_batch) return SDDC_OK;
  if (plan->live_graphs > 0) {
    char buf[160];
    snprintf(buf, sizeof(buf),
             "the plan's workspace is held by %d HIP graph(s); destroy them before reserving "
             "%d instances (reserved: %d)", plan->live_graphs, max_batch, plan->p.max_batch);
    return fail(SDDC_ERR_INVALID, buf);
  }
  SDDC_HIP(hipDeviceSynchronize());   // the old workspace may still be in use by queued kernels
  return alloc_ws(plan->p, max_batch);
}

int sddc_plan_destroy(sddc_plan* plan) {
  if (!plan) return SDDC_OK;
  std::lock_guard<std::recursive_mutex> lk(device_state_mutex());
  if (plan->live_graphs > 0)
    return fail(SDDC_ERR_INVALID, "the plan's workspace is held by live HIP graphs");
  Plan& p = plan->p;
  free_ws(p);
  void* bufs[] = {p.Dfull, p.R, p.mfma64, p.mfma32, p.vecs};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  free_chain(p);
  free_tables(p.tabM);
  free_tables(p.tabKE);
  delete plan;
  return SDDC_OK;
}

int sddc_plan_info(const sddc_plan* plan, long long* workspace_bytes, int* lds_nonlinear) {
  if (!plan) return fail(SDDC_ERR_INVALID, "null plan");
  if (workspace_bytes) *workspace_bytes = (long long)plan->p.ws_bytes;
  if (lds_nonlinear) *lds_nonlinear = nonlinear_lds_bytes(plan->p, plan->p.precision, 0);
  return SDDC_OK;
}

int sddc_transform(int kind, const double* in, double* out, int rows, int len_in, int n_out,
                   void* stream) {
  if (kind < SDDC_IDCT || kind > SDDC_DST) return fail(SDDC_ERR_INVALID, "unknown transform kind");
  if (rows < 0 || len_in < 1 || n_out < 1) return fail(SDDC_ERR_INVALID, "bad transform size");
  const bool inverse = (kind == SDDC_IDCT || kind == SDDC_IDST);
  if (kind == SDDC_IDST && n_out < len_in)
    return fail(SDDC_ERR_UNSUPPORTED, "IDST with n below the input length is not supported");
  const int L = inverse ? n_out : len_in;
  static std::mutex mu;
  static std::map<std::pair<int, int>, TableSet> cache;
  int dev = 0;
  SDDC_HIP(hipGetDevice(&dev));
  TableSet* t = nullptr;
  {
    std::lock_guard<std::recursive_mutex> dlk(device_state_mutex());
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(dev, L);
    auto it = cache.find(key);
    if (it == cache.end()) {
      TableSet ts;
      int rc = build_tables(ts, L);
      if (rc) return rc;
      it = cache.emplace(key, ts).first;
    }
    t = &it->second;
  }
  return launch_transform(*t, kind, rows, len_in, n_out, in, out, (hipStream_t)stream);
}

static int prep_input(Plan& p, int B, const double* X, const double** use, hipStream_t s,
                      double* scratch) {
  *use = X;
  if (p.symmetric) {
    int rc = launch_mask(p, B, X, scratch, s);
    if (rc) return rc;
    *use = scratch;
  }
  return SDDC_OK;
}

int sddc_theta_ops(sddc_plan* plan, int B, const double* X, double* out, void* stream) {
  if (!plan) return fail(SDDC_ERR_INVALID, "null plan");
  Plan& p = plan->p;
  int rc = check_batch(p, B);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const double* x;
  if ((rc = prep_input(p, B, X, &x, s, p.Xa))) return rc;
  if ((rc = launch_suffix(p, B, x, p.P, s))) return rc;
  return launch_spectral(p, B, x, p.P, nullptr, nullptr, 0, nullptr, nullptr, SDDC_FP64, out,
                         nullptr, s);
}

int sddc_derivatives(sddc_plan* plan, int B, const double* X, double* out, void* stream) {
  if (!plan) return fail(SDDC_ERR_INVALID, "null plan");
  Plan& p = plan->p;
  int rc = check_batch(p, B);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const double* x;
  if ((rc = prep_input(p, B, X, &x, s, p.Xa))) return rc;
  if ((rc = launch_suffix(p, B, x, p.P, s))) return rc;
  return launch_spectral(p, B, x, p.P, nullptr, nullptr, 0, nullptr, nullptr, SDDC_FP64, nullptr,
                         out, s);
}

int sddc_nlin_fx(sddc_plan* plan, int B, const double* X, double* out, void* stream) {
  if (!plan) return fail(SDDC_ERR_INVALID, "null plan");
  Plan& p = plan->p;
  int rc = check_batch(p, B);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const double* x;
  if ((rc = prep_input(p, B, X, &x, s, p.Xa))) return rc;
  if ((rc = launch_suffix(p, B, x, p.P, s))) return rc;
  if ((rc = launch_spectral(p, B, x, p.P, nullptr, nullptr, 0, nullptr, p.slabs, p.precision,
                            nullptr, nullptr, s)))
    return rc;
  if ((rc = launch_nonlinear(p, B, p.slabs, nullptr, p.F, s))) return rc;
  if ((rc = launch_rhs_scaled(p, B, p.F, nullptr, 1.0, out, s))) return rc;
  if (p.symmetric) return launch_mask(p, B, out, out, s);
  return SDDC_OK;
}

static int dfx_pipeline(Plan& p, int B, const double* dv, const double* X, const double* Ra,
                        const double* Ras, bool linear_terms, double scale, double* out,
                        hipStream_t s) {
  int rc;
  const double *x = X, *v = dv;
  if (p.symmetric) {
    if ((rc = launch_mask(p, B, X, p.Xa, s))) return rc;
    if ((rc = launch_mask(p, B, dv, p.Xb, s))) return rc;
    x = p.Xa;
    v = p.Xb;
  }
  if ((rc = launch_suffix(p, B, x, p.P, s))) return rc;
  if ((rc = launch_spectral(p, B, x, p.P, nullptr, nullptr, 0, nullptr, p.slabs, p.precision,
                            nullptr, nullptr, s)))
    return rc;
  if ((rc = launch_suffix(p, B, v, p.Pa, s))) return rc;
  // the second slab set reuses slabs2 in the nonlinear pipeline precision
  if ((rc = launch_spectral(p, B, v, p.Pa, Ra, Ras, linear_terms ? 1 : 0, p.L, p.slabs2,
                            p.precision, nullptr, nullptr, s)))
    return rc;
  if ((rc = launch_nonlinear(p, B, p.slabs, p.slabs2, p.F, s))) return rc;
  return launch_rhs_scaled(p, B, p.F, linear_terms ? p.L : nullptr, scale, out, s);
}

int sddc_nlin_dfx(sddc_plan* plan, int B, const double* dv, const double* X, double* out,
                  void* stream) {
  if (!plan) return fail(SDDC_ERR_INVALID, "null plan");
  Plan& p = plan->p;
  int rc = check_batch(p, B);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = dfx_pipeline(p, B, dv, X, nullptr, nullptr, false, 1.0, out, s))) return rc;
  if (p.symmetric) return launch_mask(p, B, out, out, s);
  return SDDC_OK;
}

int sddc_bsub(sddc_plan* plan, int B, int field, const double* g, double* f, void* stream) {
  if (!plan) return fail(SDDC_ERR_INVALID, "null plan");
  Plan& p = plan->p;
  int rc = check_batch(p, B);
  if (rc) return rc;
  if (field < 0 || field > 2) return fail(SDDC_ERR_INVALID, "field must be 0 (psi), 1 (T) or 2 (S)");
  const long long N = (long long)p.K * p.nr;
  const long long off[3] = {0, 0, 0};
  return launch_solve_ex(p, B, g, N, off, f, N, off, nullptr, 1 << field, (hipStream_t)stream);
}

static int step_once(Plan& p, int B, const double* Ra, const double* Ras, const double* src,
                     double* dst, int linear, hipStream_t s) {
  int rc;
  if ((rc = launch_spectral(p, B, src, p.P, Ra, Ras, 1, p.L, linear ? nullptr : p.slabs,
                            p.precision, nullptr, nullptr, s)))
    return rc;
  if (!linear && (rc = launch_nonlinear(p, B, p.slabs, nullptr, p.F, s))) return rc;
  if ((rc = launch_rhs(p, B, p.F, p.L, linear ? 0 : 1, p.G, s))) return rc;
  return launch_solve(p, B, p.G, dst, p.P, 7, 0, s);
}

static int run_steps(Plan& p, int B, const double* Ra, const double* Ra_s, const double* X_in,
                     double* X_out, int nsteps, int linear, hipStream_t s) {
  int rc;
  if (nsteps == 0) {
    if (X_out != X_in)
      SDDC_HIP(hipMemcpyAsync(X_out, X_in, (size_t)B * 3 * p.K * p.nr * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
    return SDDC_OK;
  }
  const double* src;
  if ((rc = prep_input(p, B, X_in, &src, s, p.Xa))) return rc;
  if ((rc = launch_suffix(p, B, src, p.P, s))) return rc;
  for (int k = 0; k < nsteps; ++k) {
    if ((rc = step_once(p, B, Ra, Ra_s, src, X_out, linear, s))) return rc;
    src = X_out;
  }
  return SDDC_OK;
}

// The plan restricted to instances [b0, b0 + nb): operators shared,
// per-instance workspace offset by b0.  b0 is a multiple of 16, so the solve's
// segment exit states stay aligned to whole workgroups.  Concurrent parts
// share only the solve's sink (stores of padded rows, never read for real
// instances).
static Plan sub_plan(const Plan& p, int b0, int nb) {
  Plan v = p;
  const size_t N = (size_t)p.K * p.nr;
  const size_t rs = p.precision == SDDC_MIXED ? sizeof(float) : sizeof(double);
  v.P = p.P + b0 * N;
  v.Pa = p.Pa + b0 * N;
  v.L = p.L + b0 * 3 * N;
  v.G = p.G + b0 * 3 * N;
  v.Xa = p.Xa + b0 * 3 * N;
  v.Xb = p.Xb + b0 * 3 * N;
  v.slabs = static_cast<char*>(p.slabs) + b0 * 7 * N * rs;
  v.slabs2 = static_cast<char*>(p.slabs2) + b0 * 7 * N * sizeof(double);
  v.F = static_cast<char*>(p.F) + b0 * 4 * N * rs;
  v.rowval = p.rowval + (size_t)b0 * p.nr;
  v.chain_EX = p.chain_EX + (size_t)b0 * 6 * p.nseg * 2 * p.NRP;
  v.ex_batch = p.ex_batch - b0;
  v.max_batch = nb;
  return v;
}

// ---------------------------------------------------------------------------
// Allocation-free step (sddc_step_static): the step's intermediates in one
// static device array per GPU instead of the plan's workspace.  Every buffer
// carved from it is written before it is read within a call (suffix sums,
// linear part, slabs, products, right-hand side, segment exit states; the
// solve's sink is only written), so nothing in it has to start at zero and
// a call does not depend on what an earlier call left there.
// ---------------------------------------------------------------------------
constexpr size_t kStaticScratchBytes = (size_t)SDDC_STATIC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

static std::mutex g_scratch_mu;
static std::vector<std::pair<int, unsigned char*>> g_scratch_ptr;   // (device, address)

// The address of this device's copy of `scratch`, looked up once per device
// with that device current (hipGetSymbolAddress); the array's name is a host
// stand-in and is never passed to a kernel or copy.
static int static_scratch(int dev, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  for (auto& e : g_scratch_ptr)
    if (e.first == dev) {
      *out = e.second;
      return SDDC_OK;
    }
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return hip_check(e != hipSuccess ? e : hipErrorUnknown, "hipGetSymbolAddress(scratch)");
  g_scratch_ptr.push_back({dev, static_cast<unsigned char*>(ptr)});
  *out = static_cast<unsigned char*>(ptr);
  return SDDC_OK;
}

// The plan restricted to B instances with its step workspace carved from
// `base` (256-byte aligned pieces); *need = the bytes the batch uses.
static Plan static_view(const Plan& p, int B, unsigned char* base, size_t* need) {
  Plan v = p;
  const size_t N = (size_t)p.K * p.nr;
  const size_t rs = p.precision == SDDC_MIXED ? sizeof(float) : sizeof(double);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* q = base + off;
    off += (bytes + 255) & ~(size_t)255;
    return q;
  };
  v.P = reinterpret_cast<double*>(take((size_t)B * N * 8));
  v.L = reinterpret_cast<double*>(take((size_t)B * 3 * N * 8));
  v.G = reinterpret_cast<double*>(take((size_t)B * 3 * N * 8));
  v.Xa = p.symmetric ? reinterpret_cast<double*>(take((size_t)B * 3 * N * 8)) : nullptr;
  v.slabs = take((size_t)B * 7 * N * 8);   // sized as the plan's workspace sizes it
  v.F = take((size_t)B * 4 * N * rs);
  v.solve_scratch = reinterpret_cast<double*>(take((4 * N + 256) * 8));
  const int exb = ((B + 3) / 4) * 4;
  v.chain_EX = reinterpret_cast<double*>(take((size_t)exb * 6 * p.nseg * 2 * p.NRP * 8));
  v.ex_batch = exb;
  v.Xb = v.Pa = nullptr;
  v.slabs2 = nullptr;
  v.rowval = v.params = nullptr;
  v.max_batch = B;
  *need = off;
  return v;
}

// Hardware queues the HIP runtime gives this process (GPU_MAX_HW_QUEUES,
// 4 when unset).
static int hw_queue_count() {
  const char* e = getenv("GPU_MAX_HW_QUEUES");
  const int n = e ? atoi(e) : 0;
  return n >= 1 ? n : 4;
}

// Parts [b0, b0 + nb) of a batch of B for `parts` forked streams: sizes are
// multiples of 16 (the last part takes the rest); fewer parts when B is small.
static std::vector<std::pair<int, int>> split_batch(int B, int parts) {
  std::vector<std::pair<int, int>> out;
  if (parts < 2 || B < 32) {
    out.push_back({0, B});
    return out;
  }
  int q = (B + parts - 1) / parts;
  q = ((q + 15) / 16) * 16;
  for (int b0 = 0; b0 < B; b0 += q) out.push_back({b0, std::min(q, B - b0)});
  return out;
}

int sddc_step(sddc_plan* plan, int B, const double* Ra, const double* Ra_s, const double* X_in,
              double* X_out, int nsteps, int linear, void* stream) {
  if (!plan) return fail(SDDC_ERR_INVALID, "null plan");
  Plan& p = plan->p;
  int rc = check_batch(p, B);
  if (rc) return rc;
  if (nsteps < 0) return fail(SDDC_ERR_INVALID, "negative step count");
  return run_steps(p, B, Ra, Ra_s, X_in, X_out, nsteps, linear, (hipStream_t)stream);
}

int sddc_step_static(sddc_plan* plan, int B, const double* Ra, const double* Ra_s,
                     const double* X_in, double* X_out, int nsteps, int linear, void* stream) {
  if (!plan) return fail(SDDC_ERR_INVALID, "null plan");
  Plan& p = plan->p;
  if (B < 0) return fail(SDDC_ERR_INVALID, "negative batch");
  if (nsteps < 0) return fail(SDDC_ERR_INVALID, "negative step count");
  int dev = 0;
  SDDC_HIP(hipGetDevice(&dev));
  if (dev != p.device) return fail(SDDC_ERR_INVALID, "the plan was created on another device");
  unsigned char* base = nullptr;
  if (int rc = static_scratch(dev, &base)) return rc;
  size_t need = 0;
  Plan v = static_view(p, B, base, &need);
  if (need > kStaticScratchBytes) {
    size_t one = 0;
    (void)static_view(p, 1, base, &one);
    char buf[200];
    snprintf(buf, sizeof(buf),
             "batch of %d needs %zu bytes of static scratch, above its %zu (about %zu instances "
             "of this size class fit)",
             B, need, kStaticScratchBytes, one ? kStaticScratchBytes / one : (size_t)0);
    return fail(SDDC_ERR_INVALID, buf);
  }
  return run_steps(v, B, Ra, Ra_s, X_in, X_out, nsteps, linear, (hipStream_t)stream);
}

int sddc_step_linearized(sddc_plan* plan, int B, const double* Ra, const double* Ra_s,
                         const double* dv, const double* X, double* out, void* stream) {
  if (!plan) return fail(SDDC_ERR_INVALID, "null plan");
  Plan& p = plan->p;
  int rc = check_batch(p, B);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = dfx_pipeline(p, B, dv, X, Ra, Ra_s, true, -1 * p.dt, p.G, s))) return rc;
  return launch_solve(p, B, p.G, out, nullptr, 7, 0, s);
}

int sddc_diagnostics(sddc_plan* plan, int B, const double* X, double* out, void* stream) {
  if (!plan) return fail(SDDC_ERR_INVALID, "null plan");
  Plan& p = plan->p;
  int rc = check_batch(p, B);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = launch_suffix(p, B, X, p.Pa, s))) return rc;
  if ((rc = launch_spectral(p, B, X, p.Pa, nullptr, nullptr, 0, nullptr, p.slabs2, SDDC_FP64,
                            nullptr, nullptr, s)))
    return rc;
  if ((rc = launch_ke_rows(p, B, p.slabs2, p.rowval, s))) return rc;
  return launch_diag_final(p, B, X, p.rowval, out, s);
}

int sddc_run_stage(sddc_plan* plan, int B, int stage, const double* X, const double* Ra,
                   const double* Ra_s, void* stream) {
  if (!plan) return fail(SDDC_ERR_INVALID, "null plan");
  Plan& p = plan->p;
  int rc = check_batch(p, B);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  switch (stage) {
    case 0: return launch_suffix(p, B, X, p.P, s);
    case 1:
      return launch_spectral(p, B, X, p.P, Ra, Ra_s, 1, p.L, p.slabs, p.precision, nullptr,
                             nullptr, s);
    case 2: return launch_nonlinear(p, B, p.slabs, nullptr, p.F, s);
    case 3: return launch_rhs(p, B, p.F, p.L, 1, p.G, s);
    case 4: return launch_solve(p, B, p.G, p.Xb, p.Pa, 7, 0, s);
    default: return fail(SDDC_ERR_INVALID, "stage must be 0..4");
  }
}

int sddc_graph_create(sddc_plan* plan, int B, const double* Ra, const double* Ra_s, double* X,
                      int nsteps, void* stream, sddc_graph** out) {
  if (!plan || !out) return fail(SDDC_ERR_INVALID, "null argument");
  Plan& p = plan->p;
  int rc = check_batch(p, B);
  if (rc) return rc;
  if (nsteps < 1) return fail(SDDC_ERR_INVALID, "nsteps must be positive");
  (void)p;
  sddc_plan* const plans[1] = {plan};
  const int Bs[1] = {B};
  const double* const ras[1] = {Ra};
  const double* const rass[1] = {Ra_s};
  double* const xs[1] = {X};
  return sddc_graph_create_multi(1, plans, Bs, ras, rass, xs, nsteps, stream, out);
}

int sddc_graph_create_multi(int nplans, sddc_plan* const* plans, const int* B,
                            const double* const* Ra, const double* const* Ra_s,
                            double* const* X, int nsteps, void* stream, sddc_graph** out) {
  if (!plans || !B || !Ra || !Ra_s || !X || !out || nplans < 1)
    return fail(SDDC_ERR_INVALID, "null argument or no plans");
  if (nsteps < 1) return fail(SDDC_ERR_INVALID, "nsteps must be positive");
  for (int c = 0; c < nplans; ++c) {
    if (!plans[c]) return fail(SDDC_ERR_INVALID, "null plan");
    if (int rc = check_batch(plans[c]->p, B[c])) return rc;
    // the eigenbasis T/S solve has not been validated under graph capture
    if (plans[c]->p.nab2_modal)
      return fail(SDDC_ERR_UNSUPPORTED, "graphs of plans with nab2_solve 2 are not supported");
  }
  // no allocation, free or device synchronisation of this library may run
  // on another thread between the capture's begin and end
  std::lock_guard<std::recursive_mutex> lk(device_state_mutex());
  sddc_graph* g = new sddc_graph();
  auto bail = [&](int rc) {
    sddc_graph_destroy(g);
    return rc;
  };
  hipStream_t s = (hipStream_t)stream;
  if (!s) {
    if (hipStreamCreateWithFlags(&g->own, hipStreamNonBlocking) != hipSuccess)
      return bail(fail(SDDC_ERR_HIP, "stream creation failed"));
    s = g->own;
  }
  // the parts: each class split into its plan's sub_batches
  struct Part {
    Plan view;
    int c, b0, nb;
  };
  std::vector<Part> parts;
  for (int c = 0; c < nplans; ++c)
    for (auto [b0, nb] : split_batch(B[c], plans[c]->p.tune_sub_batches))
      parts.push_back({sub_plan(plans[c]->p, b0, nb), c, b0, nb});
  const int np = (int)parts.size();
  auto state_of = [&](const Part& pt) {
    return X[pt.c] + pt.b0 * (size_t)3 * pt.view.K * pt.view.nr;
  };
  // one eager step per class so that every lazily set kernel attribute is in place
  int rc = SDDC_OK;
  for (int c = 0; c < nplans && !rc; ++c)
    rc = sddc_step(plans[c], B[c], Ra[c], Ra_s[c], X[c], plans[c]->p.Xb, 1, 0, s);
  if (!rc) rc = hip_check(hipStreamSynchronize(s), "warm-up");
  if (rc) return bail(rc);
  // Lanes.  Kernels on one hardware queue run one after another, and a
  // process has few queues, so with more parts than queues the parts that
  // share a queue add up, however the runtime places the graph's branches.
  // With more parts than queues, each part is timed alone (one step, written
  // to the plan's spare state rows; the second of two runs, so that no
  // first launch is counted) and the parts go onto as many lanes as there
  // are queues, longest first onto the least loaded; each lane is one forked
  // capture stream that runs its parts in turn.
  const int nl = std::min(np, hw_queue_count());
  std::vector<int> lane(np);
  for (int q = 0; q < np; ++q) lane[q] = q;
  if (nl < np) {
    std::vector<float> cost(np, 0.f);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    rc = hip_check(hipEventCreate(&t0), "event creation");
    if (!rc) rc = hip_check(hipEventCreate(&t1), "event creation");
    for (int q = 0; q < np && !rc; ++q) {
      Part& pt = parts[q];
      for (int rep = 0; rep < 2 && !rc; ++rep) {
        rc = hip_check(hipEventRecord(t0, s), "timing");
        if (!rc)
          rc = run_steps(pt.view, pt.nb, Ra[pt.c] + pt.b0, Ra_s[pt.c] + pt.b0, state_of(pt),
                         pt.view.Xb, 1, 0, s);
        if (!rc) rc = hip_check(hipEventRecord(t1, s), "timing");
        if (!rc) rc = hip_check(hipEventSynchronize(t1), "timing");
        if (!rc) rc = hip_check(hipEventElapsedTime(&cost[q], t0, t1), "timing");
      }
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (rc) return bail(rc);
    std::vector<int> order(np);
    for (int q = 0; q < np; ++q) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<float> load(nl, 0.f);
    for (int q : order) {
      const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      lane[q] = l;
      load[l] += cost[q];
    }
  }
  g->side.assign(nl, nullptr);
  g->events.assign(nl + 1, nullptr);
  for (int l = 0; l < nl; ++l)
    if (hipStreamCreateWithFlags(&g->side[l], hipStreamNonBlocking) != hipSuccess)
      return bail(fail(SDDC_ERR_HIP, "stream creation failed"));
  for (auto& e : g->events)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return bail(fail(SDDC_ERR_HIP, "event creation failed"));
  // fork: every lane's stream waits for the capture origin, runs its parts'
  // steps, and the origin waits for all of them
  if ((rc = hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "capture")))
    return bail(rc);
  int rc2 = hip_check(hipEventRecord(g->events[nl], s), "fork");
  for (int l = 0; l < nl && !rc2; ++l) {
    rc2 = hip_check(hipStreamWaitEvent(g->side[l], g->events[nl], 0), "fork wait");
    for (int q = 0; q < np && !rc2; ++q) {
      if (lane[q] != l) continue;
      Part& pt = parts[q];
      double* xq = state_of(pt);
      rc2 = run_steps(pt.view, pt.nb, Ra[pt.c] + pt.b0, Ra_s[pt.c] + pt.b0, xq, xq, nsteps, 0,
                      g->side[l]);
    }
    if (!rc2) rc2 = hip_check(hipEventRecord(g->events[l], g->side[l]), "join");
    if (!rc2) rc2 = hip_check(hipStreamWaitEvent(s, g->events[l], 0), "join wait");
  }
  hipError_t e = hipStreamEndCapture(s, &g->graph);
  rc = rc2 ? rc2 : hip_check(e, "end capture");
  // a failed capture leaves its error as the thread's last error, which the
  // next launch check of this thread would report: clear it here
  if (rc) (void)hipGetLastError();
  if (!rc) rc = hip_check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "instantiate");
  if (rc) return bail(rc);
  for (int c = 0; c < nplans; ++c) {
    bool seen = false;
    for (sddc_plan* q : g->plans) seen = seen || q == plans[c];
    if (!seen) {
      g->plans.push_back(plans[c]);
      ++plans[c]->live_graphs;
    }
  }
  *out = g;
  return SDDC_OK;
}

int sddc_graph_launch(sddc_graph* g, void* stream) {
  if (!g) return fail(SDDC_ERR_INVALID, "null graph");
  hipStream_t s = stream ? (hipStream_t)stream : g->own;
  return hip_check(hipGraphLaunch(g->exec, s), "graph launch");
}

int sddc_graph_destroy(sddc_graph* g) {
  if (!g) return SDDC_OK;
  std::lock_guard<std::recursive_mutex> lk(device_state_mutex());
  for (sddc_plan* q : g->plans) --q->live_graphs;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->own) (void)hipStreamDestroy(g->own);
  for (hipStream_t st : g->side)
    if (st) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : g->events)
    if (ev) (void)hipEventDestroy(ev);
  delete g;
  return SDDC_OK;
}

}  // extern "C"
