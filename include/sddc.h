/*
 * libsddc -- C ABI of the MI355X (gfx950) pseudospectral double-diffusive
 * convection time step.
 *
 * All compute entry points take device pointers and a HIP stream (passed as
 * void*; NULL = the default stream) and are stream-ordered.  They return an
 * int status (SDDC_OK = 0); sddc_last_error() gives the message of the last
 * failure on the calling thread.  No entry point allocates device memory
 * except sddc_plan_create / sddc_plan_reserve, sddc_krylov_create,
 * sddc_newton_batch_create and the first call of a transform at a new
 * length (its twiddle tables are cached).
 *
 * Threads: one device-state lock (process-wide) is held for the whole of a
 * graph capture (sddc_graph_create*) and by every call that allocates,
 * frees, copies synchronously or synchronises the device (plan, Krylov and
 * batched-Newton create/reserve/destroy, graph destroy, new transform
 * lengths), so such a call on another thread waits for a capture to finish
 * instead of invalidating it.  Compute calls take no lock, and on MI355X
 * kernel launches of another thread (on any stream) during a capture can
 * invalidate it: callers keep compute calls of other threads off the device
 * while sddc_graph_create* runs.  Calls on one plan (or one Krylov object)
 * must not overlap in time on different threads.
 *
 * Batches: a plan describes one size class (N_fm, N_r, d, dt, Pr, Tau); B
 * instances of that class are stored back to back, each in the reference's
 * state layout X[f*N + k*nr + i] (f = psi,T,S; N = N_fm*nr; nr = N_r-1).
 * Ra and Ra_s are per-instance device arrays of length B.
 *
 * Reference interfaces are cited as file:line of
 * mannixp/SpectralDoubleDiffusiveConvection.
 */
#ifndef SDDC_H
#define SDDC_H

#ifdef __cplusplus
extern "C" {
#endif

enum sddc_status {
  SDDC_OK = 0,
  SDDC_ERR_INVALID = 1,      /* bad argument (maps to ValueError)          */
  SDDC_ERR_HIP = 2,          /* HIP runtime error                          */
  SDDC_ERR_UNSUPPORTED = 3,  /* size or mode the library does not support  */
  SDDC_ERR_ODD_NFM = 4       /* odd N_fm (Matrix_Operators.py:758-759)     */
};

enum sddc_precision {
  SDDC_FP64 = 0,   /* plain method: everything in float64                  */
  SDDC_MIXED = 1   /* float32 transforms, products and Chebyshev GEMMs of  */
                   /* the nonlinear term; float64 linear terms and solves   */
};

enum sddc_transform_kind { SDDC_IDCT = 0, SDDC_IDST = 1, SDDC_DCT = 2, SDDC_DST = 3 };

/* Host-side operator set, built exactly as Build_Matrix_Operators
 * (Main.py:179-225) builds it.  Matrices are row-major float64. */
typedef struct sddc_operators {
  int N_fm;                 /* theta modes K (even)                          */
  int N_r;                  /* Chebyshev degree; nr = N_r - 1 unknowns        */
  double d, dt, Pr, Tau;
  const double* D;          /* (N_r+1)^2  cheb_radial D (Matrix_Operators.py:10-28) */
  const double* R;          /* N_r+1      radial points                       */
  const double* Dr;         /* nr^2  D[1:-1,1:-1] (Matrix_Operators.py:764)   */
  const double* D2s;        /* nr^2  (D@D)[1:-1,1:-1] (A2_SINE, :217)         */
  const double* D2r2;       /* nr^2  (diag(1/R^2) D@D)[1:-1,1:-1] (A2_SINE_R2, :498) */
  const double* IR2v;       /* nr    1/r^2 (A2_SINE, :216)                    */
  const double* IR4v;       /* nr    1/r^4 (A2_SINE_R2, :496)                 */
  const double* DT0;        /* nr    A_T/r^2 (Main.py:201-202)                */
  const double* Rsq;        /* nr    r^2 (R2, Matrix_Operators.py:79-94)      */
  const double* GR;         /* nr    R_1^2/r^2 (kGR_RT, :113)                 */
  const double* D2_A4;      /* nr^2  args_A4 D2 (Main.py:212)                 */
  const double* IR2_A4;     /* nr    diagonal of args_A4 IR2 (Main.py:213)    */
  const double* IR4_A4;     /* nr    diagonal of args_A4 IR4 (Main.py:214)    */
  const double* Linv_A4;    /* N_fm*nr^2  A4_TSTEP_MATS(Pr*dt) (Matrix_Operators.py:1089-1112) */
  const double* Linv_T;     /* N_fm*nr^2  NAB2_TSTEP_MATS(dt)  (Matrix_Operators.py:1014-1030) */
  const double* Linv_S;     /* N_fm*nr^2  NAB2_TSTEP_MATS(Tau*dt)                              */
  /* Optional (NULL: the NAB2 solves use the per-mode blocks above).  The
   * eigenbasis of P = r^2 - h r^2 Lap, whose shifts P + h j(j+1) I are the
   * NAB2 matrices of mode j (h = dt for T, Tau*dt for S): nr eigenvalues lam,
   * then V and V^-1 (nr x nr, row-major) with P = V diag(lam) V^-1.  With
   * them each NAB2 chain decouples into nr scalar recurrences. */
  const double* modes_T;    /* nr + 2 nr^2, h = dt      */
  const double* modes_S;    /* nr + 2 nr^2, h = Tau*dt  */
} sddc_operators;

typedef struct sddc_plan sddc_plan;

/* Launch-shape choices of a plan (no counterpart in the reference; the
 * autotuner varies them).  0 leaves a field to the library's default.
 *   chain_segment    rows per segment of the parity chains of the solves
 *                    (>= rows per chain: one segment, no join)
 *   solve_instances  instances per solve workgroup: 1, 2 or 4 (quad-lane
 *                    kernels), 3 (quad-lane, 2 per A4 workgroup and 4 per
 *                    NAB2 workgroup), 5 (as 3, the A4 chains with two
 *                    exchanges per row and one streamed block), 6 (as 5,
 *                    3 per A4 workgroup), 7 and 8 (as 5 with four and six
 *                    rows of operator data in flight), 9 (one instance per
 *                    A4 workgroup, two exchanges, six rows in flight), 10
 *                    (one instance per A4 workgroup, one exchange, four
 *                    rows in flight), 16 (MFMA kernel).  1-4 and 10 give the
 *                    same bits, as do 5-9; these two groups and 16 sum in
 *                    different orders (equal to rounding), so the choice is
 *                    fixed per plan
 *   sub_batches      graphs only: the batch is split into this many parts
 *                    (1-8), stepped on forked streams inside the graph (see
 *                    sddc_graph_create_multi), so that the latency-bound
 *                    solves of one part overlap the transforms of another.
 *                    Results are the same bits as one batch.
 *   nab2_solve       the T and S solves: 0 or 1 chains of per-mode block
 *                    inverses (the reference's form), 2 scalar recurrences
 *                    in the eigenbasis of the operator set's modes_T /
 *                    modes_S (equal to rounding, about 1e-14 relative; a
 *                    separate launch after the A4 chains, about three times
 *                    faster than the block chains on its own).
 *   spectral_instances  instances per workgroup of the spectral stage (1-4;
 *                    0: from the number of (mode tile, instance) pairs).
 *                    The same bits whatever the value. */
typedef struct sddc_tuning {
  int chain_segment;
  int solve_instances;
  int sub_batches;
  int nab2_solve;
  int spectral_instances;
  int reserved[3];
} sddc_tuning;

/* Plan lifecycle.  Replaces the operator tuple of Build_Matrix_Operators
 * (Main.py:179-225): uploads the operators once and sizes workspace for
 * max_batch instances.  symmetric=1 applies the equatorial mask of Eq_SYM
 * (Main.py:137-176) to inputs and outputs. */
int sddc_plan_create(const sddc_operators* ops, int precision, int symmetric,
                     int max_batch, sddc_plan** out);
/* sddc_plan_create with explicit launch shapes (tuning may be NULL). */
int sddc_plan_create_tuned(const sddc_operators* ops, int precision, int symmetric,
                           int max_batch, const sddc_tuning* tuning, sddc_plan** out);
/* The launch shapes a plan uses (defaults resolved, except two chosen per
 * call: solve_instances 0 is 1 below 8 instances, 2 below 64 and 3 from 64
 * on, all of one bit group, so a plan's results do not depend on the
 * batch size of a call; spectral_instances 0 follows the call's number of
 * (mode tile, instance) pairs, any value giving the same bits). */
int sddc_plan_tuning(const sddc_plan* plan, sddc_tuning* out);
/* Grow the plan's workspace to max_batch instances (no-op if it holds that
 * many).  Growing frees the old workspace, so it is refused with
 * SDDC_ERR_INVALID while a graph created on the plan is alive; so is
 * sddc_plan_destroy.  Destroy the graphs first. */
int sddc_plan_reserve(sddc_plan* plan, int max_batch);
int sddc_plan_destroy(sddc_plan* plan);
int sddc_plan_info(const sddc_plan* plan, long long* workspace_bytes, int* lds_nonlinear);

/* Transforms.py:73-129.  rows independent rows of length len_in.
 *   SDDC_IDCT  IDCT(f_hat, n)  (Transforms.py:116-129)  out rows have n_out
 *   SDDC_IDST  IDST(g_hat, n)  (Transforms.py:87-100)   (n_out >= len_in)
 *   SDDC_DCT   DCT(f, n)       (Transforms.py:102-114)  out rows have
 *   SDDC_DST   DST(g, n)       (Transforms.py:73-85)    min(n_out, len_in) */
int sddc_transform(int kind, const double* in, double* out, int rows, int len_in,
                   int n_out, void* stream);

/* theta operators of psi (Matrix_Operators.py); out gets [B][4][N]:
 *   J_theta_RT (:436-472), DT0_theta (:131-189), A2_SINE (:192-245),
 *   A2_SINE_R2 (:475-526).  psi is the first N values of each instance. */
int sddc_theta_ops(sddc_plan* plan, int B, const double* X, double* out, void* stream);

/* Derivatives (Matrix_Operators.py:630-740): the nine (nr, N_fm) slabs
 * JT,kDpsi,komega,DT,DC,omega,Dpsi,kT,kC as [B][9][nr][N_fm] float64. */
int sddc_derivatives(sddc_plan* plan, int B, const double* X, double* out, void* stream);

/* NLIN_FX (Matrix_Operators.py:743-804): out = F(X,X), [B][3N]. */
int sddc_nlin_fx(sddc_plan* plan, int B, const double* X, double* out, void* stream);

/* NLIN_DFX (Matrix_Operators.py:807-898): out = F(X,dv) + F(dv,X). */
int sddc_nlin_dfx(sddc_plan* plan, int B, const double* dv, const double* X,
                  double* out, void* stream);

/* Back substitutions with the plan's pre-inverted blocks.
 *   field 0: A4_BSub_TSTEP_V2 (Matrix_Operators.py:1115-1194), dt = Pr*dt
 *   field 1: NAB2_BSub_TSTEP_V2 (:1033-1086) with Linv_T, dt
 *   field 2: NAB2_BSub_TSTEP_V2 with Linv_S, Tau*dt
 * g and f are [B][N] (one field per instance). */
int sddc_bsub(sddc_plan* plan, int B, int field, const double* g, double* f, void* stream);

/* nsteps applications of Step_Python (Main.py:255-283) with the loop's
 * X = X_SYM*X_new (Main.py:323).  X_in and X_out are [B][3N]; they may
 * alias.  linear=1 drops the nonlinear term (Main.py:264). */
int sddc_step(sddc_plan* plan, int B, const double* Ra, const double* Ra_s,
              const double* X_in, double* X_out, int nsteps, int linear, void* stream);

/* sddc_step for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The step's intermediates (suffix sums,
 * linear part, slabs, products, right-hand side, solve exit states) live in
 * one static device array of SDDC_STATIC_SCRATCH_BYTES per GPU, held from
 * the first plan created on that GPU until the process ends, not in the
 * plan's workspace; the call makes no hipMalloc, hipFree or other allocating
 * call, and B is not limited by the plan's reserved batch.  A batch whose
 * intermediates do not fit the array is refused with SDDC_ERR_INVALID (the
 * message gives how many instances of the class fit: 25 at N_r = 40,
 * N_fm = 512 in mixed precision, 23 in float64, with unsegmented chains).
 * Every call on one GPU uses
 * the same array: calls that overlap in time on one GPU (other streams,
 * other threads, other plans) race.  Results are the same bits as
 * sddc_step. */
#define SDDC_STATIC_SCRATCH_BYTES (64ll << 20)
int sddc_step_static(sddc_plan* plan, int B, const double* Ra, const double* Ra_s,
                     const double* X_in, double* X_out, int nsteps, int linear, void* stream);

/* The step linearised about X applied to dv: PDFX without "- dv"
 * (Main.py:498-521, _ContinC Main.py:804-827). */
int sddc_step_linearized(sddc_plan* plan, int B, const double* Ra, const double* Ra_s,
                         const double* dv, const double* X, double* out, void* stream);

/* Diagnostics of the time loop (Main.py:292-295): out [B][6] =
 * (||X||_2, Kinetic_Energy (Main.py:71-134), Nusselt(T) and Nusselt(S) at R_1
 * (Main.py:41-68, the value Nusselt returns), Nusselt(T) and Nusselt(S) at R_2
 * (the other end of NuM1, which Nusselt's check compares, Main.py:63-66)). */
int sddc_diagnostics(sddc_plan* plan, int B, const double* X, double* out, void* stream);

/* Capture nsteps of sddc_step for a fixed batch (pointers fixed) into a HIP
 * graph; replay with sddc_graph_launch.  No counterpart in the reference:
 * replaces its Python time loop (Main.py:286-324) for launch-bound runs. */
typedef struct sddc_graph sddc_graph;
int sddc_graph_create(sddc_plan* plan, int B, const double* Ra, const double* Ra_s,
                      double* X, int nsteps, void* stream, sddc_graph** out);
/* One graph for several size classes: nsteps in-place steps of batch
 * X[c] (B[c] instances, parameters Ra[c], Ra_s[c]) of every plan c.  The
 * parts (each class split by its plan's sub_batches) run concurrently on
 * forked streams inside the graph, at most one per hardware queue
 * (GPU_MAX_HW_QUEUES, 4 when unset): with more parts than that, each part
 * is timed alone for one step at creation (into the plan's spare state
 * rows; X is not changed) and the parts are spread over the streams by that
 * time, longest first onto the least loaded, each stream running its parts
 * in turn.  No counterpart in the reference: batches parameter sweeps over
 * mixed resolutions into one launch. */
int sddc_graph_create_multi(int nplans, sddc_plan* const* plans, const int* B,
                            const double* const* Ra, const double* const* Ra_s,
                            double* const* X, int nsteps, void* stream, sddc_graph** out);
int sddc_graph_launch(sddc_graph* g, void* stream);
int sddc_graph_destroy(sddc_graph* g);

/* Matrix-free Newton-Krylov steady solves (_Newton, Main.py:430-554), float64
 * on the device: restarted GMRES(restart) with CGS2 Gram-Schmidt on a Krylov
 * basis held by the sddc_krylov object (its only allocation).  One instance
 * (B = 1) per call; X, b, x are device vectors of 3N values. */
typedef struct sddc_krylov sddc_krylov;
typedef struct sddc_krylov_opts {
  double tol_newton;      /* stop when ||dv|| / ||X|| <= tol_newton (Main.py:518)      */
  double tol_gmres;       /* Newton: Krylov atol = tol_gmres * ||F|| (Main.py:534)    */
  double rtol, atol;      /* sddc_gmres: stop at ||b - J x|| <= max(atol, rtol ||b||) */
  int max_restarts;       /* GMRES cycles (lgmres maxiter, Main.py:534: 250)          */
  int max_newton;         /* Newton iterations (Main.py:518: 5)                       */
  int finite_difference;  /* 1: J v by finite differences of the step, 0: PDFX        */
  int check_every;        /* sddc_newton_batch: Arnoldi steps between reads of the    */
                          /* instances' convergence flags (0: 4)                      */
  int reserved[4];
} sddc_krylov_opts;
typedef struct sddc_krylov_info {
  int converged, iterations, restarts;
  double residual, rhs_norm;
} sddc_krylov_info;
typedef struct sddc_newton_info {
  int converged, iterations, krylov_iterations;
  double error;
} sddc_newton_info;
int sddc_krylov_create(sddc_plan* plan, int restart, sddc_krylov** out);
int sddc_krylov_destroy(sddc_krylov* k);
/* Solve J x = b, J = PDFX at X (Step'(X) - I); x holds the initial guess. */
int sddc_gmres(sddc_krylov* k, const double* Ra, const double* Ra_s, const double* X,
               const double* b, double* x, const sddc_krylov_opts* opts, sddc_krylov_info* info,
               void* stream);
/* Newton iteration on Step(X) - X = 0 in place (X is overwritten). */
int sddc_newton(sddc_krylov* k, const double* Ra, const double* Ra_s, double* X,
                const sddc_krylov_opts* opts, sddc_newton_info* info, void* stream);

/* Batched Newton-Krylov: _Newton (Main.py:430-554) for B <= max_batch
 * instances of the plan's size class at once (Ra, Ra_s: device arrays of B
 * values; X: [B][3N], overwritten; info: B host structs).  Each instance runs
 * sddc_newton's algorithm (tol_newton, tol_gmres, max_newton, max_restarts,
 * finite_difference of opts), in lockstep with the others: one J v is one
 * batched launch of the step kernels, and the Hessenberg least-squares
 * update (Givens rotations, residual estimate, back substitution) runs on
 * the device, so the host reads B flags only every opts->check_every
 * Arnoldi steps, per restart and per Newton iteration.  Converged instances
 * are frozen while the others iterate, so each instance's result is the same
 * bits as the same call with that instance alone (B = 1).  The object holds
 * (restart + 1) * max_batch * 3N doubles of basis and reserves the plan's
 * workspace for max_batch instances. */
typedef struct sddc_newton_batch sddc_newton_batch;
int sddc_newton_batch_create(sddc_plan* plan, int restart, int max_batch,
                             sddc_newton_batch** out);
int sddc_newton_batch_destroy(sddc_newton_batch* k);
int sddc_newton_batched(sddc_newton_batch* k, int B, const double* Ra, const double* Ra_s,
                      double* X, const sddc_krylov_opts* opts, sddc_newton_info* info,
                      void* stream);

/* One pseudo-arclength continuation step in Ra_T (_ContinC, Main.py:754-950)
 * on the device.  Y = (X, Ra_T) and Y_dot are device vectors of 3N+1 values,
 * updated in place: predictor (tangent from DF_X xi = -DF_Ra when dt < 10 or
 * Y_dot = 0, else Y_dot), Newton corrector on the bordered system with
 * delta = 1/(3N) (GMRES on the device), halving ds after 5 corrector
 * iterations, doubling it after convergence within 4, and the new unit
 * tangent.  *ds is updated.  Ra_s is a device scalar.  The Krylov options
 * give tol_newton, tol_gmres, rtol, max_restarts and finite_difference.
 * info->status reports the reference's failure cases (Main.py:859,891,923)
 * without an error code; a tangent solve that does not converge leaves
 * Y_dot = 0 and tangent_ok = 0 (Main.py:943-944). */
enum {
  SDDC_ARC_OK = 0,
  SDDC_ARC_PREDICTOR_FAILED = 1,  /* GMRES for xi did not converge (Main.py:859) */
  SDDC_ARC_STEP_TOO_SMALL = 2,    /* ds < tol_newton after halving (Main.py:891) */
  SDDC_ARC_CORRECTOR_FAILED = 3   /* GMRES for dY did not converge (Main.py:923) */
};
typedef struct sddc_arclength_info {
  int status;             /* SDDC_ARC_*                                    */
  int iterations;         /* corrector iterations of the accepted step     */
  int halvings;           /* ds halvings                                   */
  int krylov_iterations;  /* GMRES iterations, all solves                  */
  int tangent_ok;         /* 1: the tangent solve converged                */
  int reserved;
  double error_X, error_mu;   /* last corrector errors (Main.py:925-926)   */
} sddc_arclength_info;
int sddc_arclength_step(sddc_krylov* k, const double* Ra_s, double* Y, double* Y_dot,
                        double sign, double* ds, const sddc_krylov_opts* opts,
                        sddc_arclength_info* info, void* stream);

/* Launch one stage of the step on the plan's workspace (for timing and
 * profiling; no counterpart in the reference).  stage: 0 suffix sums,
 * 1 spectral stage, 2 nonlinear products, 3 right-hand side, 4 solves.
 * X is the state read by stages 0-1; stage 4 writes into plan scratch.  A
 * full sddc_step on the same batch must have run first. */
int sddc_run_stage(sddc_plan* plan, int B, int stage, const double* X, const double* Ra,
                   const double* Ra_s, void* stream);

const char* sddc_last_error(void);
const char* sddc_version(void);

#ifdef __cplusplus
}
#endif
#endif /* SDDC_H */
