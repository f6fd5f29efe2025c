/*
 * probit_amd.h -- C ABI of the MI355X (gfx950) GP-posterior library.
 *
 * The library replaces the dense linear algebra that bb515/probit runs inside
 * JAX/XLA for LaplaceGP / VBGP (probit/approximators.py).  Every entry point
 * is extern "C", takes plain pointers and sizes, returns an int status and
 * enqueues its work on the HIP stream it is given (NULL = default stream).
 *
 * Ownership: every pointer argument is caller-owned device memory unless
 * noted.  The library allocates nothing on the device; it never frees caller
 * memory.  Workspaces are caller-owned blocks sized by the *_workspace_size
 * queries; a fitted problem's factor stays resident in its workspace until
 * the caller reuses or frees it.
 *
 * Status: 0 = OK; negative = invalid argument (-i: argument i);
 * positive = numerical info (see PGP_INFO_*).  No C++ exception crosses the
 * ABI.  pgp_last_error() returns a message for the calling thread's last
 * failure.
 *
 * Layout: matrices are column-major with leading dimension ld.  Symmetric
 * N x N matrices in a workspace are padded to Np = ceil(N/64)*64 with an
 * identity block so that all tiles are full 64 x 64 blocks.
 */
#ifndef PROBIT_AMD_H
#define PROBIT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGP_MAX_DIM 32 /* input dimension D <= 32 */
#define PGP_MAX_CUT 17 /* J + 1 cutpoints, J <= 16 classes */

/* covariance families (mlkernels EQ / Matern12, probit/approximators.py:171) */
enum { PGP_KERNEL_EQ = 0, PGP_KERNEL_MATERN12 = 1 };

/* likelihoods (probit/utilities.py:56-60 and :151-192) */
enum {
  PGP_LIK_GAUSSIAN = 0,      /* log_gaussian_likelihood */
  PGP_LIK_PROBIT = 1,        /* log_probit_likelihood, exact derivatives */
  PGP_LIK_PROBIT_CLOSED = 2  /* grad_/hessian_log_probit_likelihood (BOUNDS) */
};

/* precision of the Cholesky factor and its trailing updates */
enum {
  PGP_PREC_FP64 = 0, /* FP64 storage, FP64 MFMA updates */
  PGP_PREC_FP32 = 1, /* FP32 storage, FP32 MFMA updates, FP64 diagonal blocks */
  PGP_PREC_BF16 = 2  /* FP32 storage, BF16 MFMA updates with FP32 accumulation */
};

enum {
  PGP_INFO_NOT_PD = 1,       /* a pivot of the factor was <= 0 or NaN */
  PGP_INFO_NEGATIVE_W = 2,   /* some precision W_i < 0 (B-form undefined) */
  PGP_INFO_MAXITER = 4,      /* fixed-point loop hit maxiter before tol */
  PGP_INFO_STALLED = 8,      /* a solve's bounded wait for another workgroup
                                ran out (results invalid) */
  PGP_INFO_ESCALATED = 16    /* BF16 trailing updates lost positive
                                definiteness; the fit continued with FP32
                                updates (results valid) */
};

/* Kernel k(x, x') = variance * kappa(r^2),
 *   r^2 = sum_d q_d / lengthscale_d^2,
 *   q_d = (x_d - x'_d)^2                       if period_d == 0
 *   q_d = 4 sin^2(pi (x_d - x'_d) / period_d)  otherwise,
 *   kappa = exp(-r^2 / 2) (EQ) or exp(-r) (Matern12).
 * This is mlkernels' `variance * EQ().stretch(l).periodic(p)`. */
typedef struct pgp_kernel {
  int32_t kind;
  int32_t dim;
  int32_t reserved0;
  int32_t reserved1;
  double variance;
  double lengthscale[PGP_MAX_DIM];
  double period[PGP_MAX_DIM]; /* 0 = not periodic in that dimension */
} pgp_kernel_t;

/* Likelihood parameters: noise_std and, for probit, J+1 cutpoints with
 * cutpoints[0] = -inf and cutpoints[J] = +inf (probit/utilities.py:47-53). */
typedef struct pgp_likelihood {
  int32_t kind;
  int32_t J;
  int32_t single_precision_bounds; /* PGP_LIK_PROBIT_CLOSED: BOUNDS["single"] */
  int32_t reserved;
  double noise_std;
  double cutpoints[PGP_MAX_CUT];
} pgp_likelihood_t;

/* Options of a fixed-point solve (LaplaceGP Newton / VBGP FPI). */
typedef struct pgp_fit_options {
  double tolerance;  /* ||z_{k+1} - z_k||_2 <= tolerance stops (jaxopt rule) */
  int32_t maxiter;   /* 100 = jaxopt's default */
  int32_t precision; /* PGP_PREC_* */
  int32_t refine;    /* FP64 iterative-refinement steps per solve */
  int32_t flags;     /* PGP_FIT_* bits, 0 by default */
} pgp_fit_options_t;

/* pgp_fit_options_t.flags: run every Newton iteration's launches directly,
 * never from a captured HIP graph (the single-problem fit otherwise replays a
 * graph of the iteration once its launch structure recurs in a later fit on
 * the same workspace; the results are the same bits either way). */
#define PGP_FIT_NO_GRAPH 1

/* Counters of the Laplace fit's graph use since the library loaded: graphs
 * captured, Newton iterations replayed from a graph and iterations launched
 * directly.  No counterpart in the reference (whose iterations run under a
 * JAX trace); for callers checking that repeated fits replay graphs. */
int pgp_fit_graph_stats(int64_t* captured, int64_t* graph_iterations,
                        int64_t* direct_iterations);

/* Host-side result of a fit. */
typedef struct pgp_fit_result {
  int32_t iterations;
  int32_t status; /* OR of PGP_INFO_* */
  double step_norm;      /* last ||z_{k+1} - z_k||_2 */
  double sum_log_lik;    /* sum_i log p(y_i | f_i) at the returned weight */
  double half_f_dot_a;   /* 0.5 f^T a */
  double log_det;        /* log |B| (Laplace) or log |sigma^2 I + K| (VB) */
  double objective;      /* objective_LA / objective_VB value */
} pgp_fit_result_t;

/* A problem: training data and caller-owned outputs (all device pointers). */
typedef struct pgp_problem {
  int64_t N;        /* training points */
  int32_t D;        /* input dimension */
  int32_t reserved;
  const double* X;  /* N x D, row-major */
  const double* y;  /* N targets (Gaussian) or class indices 0..J-1 (probit) */
  double* weight;   /* N, output: a with f = K a */
  double* precision;/* N, output: W = -d2 log p(y | f) */
  double* latent;   /* N, output: f = K a (may be NULL) */
} pgp_problem_t;

/* ------------------------------------------------------------------ */
/* diagnostics                                                          */
/* ------------------------------------------------------------------ */
const char* pgp_last_error(void);
int pgp_version(void);

/* The mapping every fit entry point applies to a problem's device status
 * word (int32 at the workspace header): PGP_INFO_NOT_PD / PGP_INFO_STALLED
 * are ORed into result->status, and result->objective becomes NaN where
 * NOT_PD, STALLED or NEGATIVE_W mark the numbers invalid (the reference's
 * JAX path yields NaN there).  Host only; exported for callers that read
 * status words themselves.  No counterpart in the reference. */
int pgp_fit_result_apply_info(pgp_fit_result_t* result, int32_t info_word);

/* Opt-in launch timer for measurement: with bit c of `mask` set, launches of
 * kernel class c (0 = outer trailing SYRK update of the Cholesky, 1 = its
 * per-column steps: diagonal factor, panel, column update, 2 = TRSV,
 * 3 = Gram assembly, 4 = TRSM, 5 = the FP64 correction of the predictive
 * variance, 6 = the merged column kernel of the Cholesky, whose super-steps
 * class 1 then leaves out) are bracketed by HIP events on their stream,
 * with the algorithmic flops / bytes each launch performs.  pgp_timer_read
 * waits for the recorded events and returns the class totals.  Not
 * thread-safe; no counterpart in the reference (measurement only). */
int pgp_timer_enable(uint32_t mask);
int pgp_timer_reset(void);
int pgp_timer_read(int32_t cls, int64_t* launches, double* total_ms,
                   double* algo_flops, double* algo_bytes);

/* Schedule table of the blocked Cholesky (the autotuner's table; no
 * counterpart in the reference, whose factorisation is LAPACK's inside XLA:
 * probit/implicit/Laplace.py:24).  One entry per precision (PGP_PREC_*) and
 * size class: class 0 holds problems with padded N <= PGP_TUNE_CLASS_NP,
 * class 1 the larger ones, and every factorisation of a batch runs each
 * class with its own entry, so a problem's schedule depends only on its own
 * size.  superstep (1..16): block columns per super-step (rank-64 superstep
 * trailing updates; changes the rounding of FP32 / BF16 factors);
 * column_max: super-steps whose first column has at most this many
 * workgroups run one launch per block column (0: never); fuse_max: update
 * launches of at most this many tiles factor the next diagonal block inside
 * the launch (0: never); column_merge: such column-form super-steps run
 * one launch per block column (0), the super-step's diagonal block of tiles
 * in one launch with in-launch tile hand-offs and then the rows below it in
 * one launch (1), or every tile in one launch (2).  column_max, fuse_max and column_merge change
 * only the launches, never the results.  The table is process-wide, read at every
 * factorisation and guarded by a mutex; probit_amd.autotune fills it from
 * timed trials and persists it. */
#define PGP_TUNE_CLASSES 2
#define PGP_TUNE_CLASS_NP 1024
typedef struct pgp_tune_entry {
  int32_t superstep;
  int32_t column_max;
  int32_t fuse_max;
  int32_t column_merge;
} pgp_tune_entry_t;
int pgp_tune_get(int32_t precision, int32_t size_class, pgp_tune_entry_t* out);
int pgp_tune_set(int32_t precision, int32_t size_class, const pgp_tune_entry_t* entry);

/* ------------------------------------------------------------------ */
/* kernel-level entry points                                            */
/* ------------------------------------------------------------------ */

/* Gram / cross-Gram matrix: out[i + j*ld] = k(X_i, Z_j) for i < N, j < M.
 * dtype 0 = FP64 out, 1 = FP32 out.  If Z == X and M == N only the lower
 * tiles are evaluated and mirrored.  Replaces `B.dense(prior(theta)(X))`
 * (probit/implicit/Laplace.py:7) and `kernel(X, X_test)`
 * (probit/approximators.py:173). */
int pgp_gram(const pgp_kernel_t* kernel, int dtype, const double* X, int64_t N,
             const double* Z, int64_t M, void* out, int64_t ld, void* stream);

/* Elementwise log-likelihood and its first two derivatives in f
 * (ll, d1 = d log p / d f, d2 = d^2 log p / d f^2).  Replaces the
 * jit(vmap(...)) wrappers of probit/approximators.py:92-104 around
 * probit/utilities.py:56-70 / :151-192.  Any output may be NULL. */
int pgp_likelihood_eval(const pgp_likelihood_t* lik, const double* f,
                        const double* y, int64_t N, double* ll, double* d1,
                        double* d2, void* stream);

/* Ordinal predictive class probabilities, out[i * J + j] = P(y_i = j) =
 * Phi((b_{j+1} - mean_i) / s_i) - Phi((b_j - mean_i) / s_i),
 * s_i = sqrt(var_i + noise_std^2).  Replaces probit_predictive_distributions
 * (probit/utilities.py:232-249). */
int pgp_probit_predictive(const pgp_likelihood_t* lik, const double* mean,
                          const double* var, int64_t M, double* out, void* stream);

/* Ordinal test metrics of the reference's classification example
 * (calculate_metrics, examples/classification.py:325-341) from class
 * probabilities P (M x J, row-major: pgp_probit_predictive's output) and
 * test labels y_test (M class indices as doubles): y_pred (M int32) =
 * argmax_j P_ij (first maximum); predictive_likelihood (M x M row-major, may
 * be NULL) = P[:, y_test] -- the reference's column indexing, an M x M
 * matrix, not the per-point probabilities; metrics (5 device doubles) =
 * [incorrect count, correct count, mean absolute error, log predictive
 * probability = sum over all M x M entries of log predictive_likelihood,
 * mean zero-one error].  workspace: 3 M caller-owned doubles.  A label
 * outside 0..J-1 gives NaN entries. */
int pgp_ordinal_metrics(const double* P, int32_t J, const double* y_test, int64_t M,
                        int32_t* y_pred, double* predictive_likelihood, double* metrics,
                        double* workspace, void* stream);

/* In-place lower Cholesky A = L L^T of an N x N matrix (N % 64 == 0),
 * storage double (PGP_PREC_FP64) or float (FP32/BF16 modes).  dinv
 * (caller-owned, N * 64 doubles) receives the FP64 inverses of the 64 x 64
 * diagonal blocks of L (what pgp_trsv / pgp_trsm recompute from L); logdiag
 * (N doubles, may be NULL) receives log L_ii in FP64; info (one device int32)
 * gets PGP_INFO_NOT_PD on a non-positive pivot.  Replaces B.cholesky
 * (probit/implicit/Laplace.py:24, VB.py:10,25). */
int pgp_potrf(int precision, void* A, int64_t N, int64_t ld, double* dinv,
              double* logdiag, int32_t* info, void* stream);

/* GP prior samples f = L_K z, L_K the lower Cholesky factor of
 * K + jitter I with K = kernel(X, X) (FP64 throughout): z and f are N x S,
 * column-major (one sample per column, leading dimensions ldz, ldf >= N).
 * The workspace (pgp_prior_sample_workspace_size(N) bytes, caller-owned)
 * receives K and its factor (padded to a multiple of 64 with the identity).
 * A K + jitter I that is not positive definite gives NaN samples, as the
 * reference's jnp.linalg.cholesky does.  Replaces the sampling in the data
 * generators (examples/regression.py:52-58,
 * examples/classification.py:227-234). */
int64_t pgp_prior_sample_workspace_size(int64_t N);
int pgp_prior_sample(const pgp_kernel_t* kernel, const double* X, int64_t N,
                     double jitter, const double* z, int64_t S, int64_t ldz,
                     double* f, int64_t ldf, void* workspace,
                     int64_t workspace_bytes, void* stream);

/* Triangular solve with the lower factor: trans = 0 solves L x = b,
 * trans = 1 solves L^T x = b.  dinv (caller-owned, N * 64 doubles) receives
 * the FP64 inverses of the 64 x 64 diagonal blocks; b (N doubles) is
 * overwritten with scratch and x receives the solution (N % 64 == 0).
 * Replaces the two solves of B.cholesky_solve (probit/implicit/VB.py:11). */
int pgp_trsv(int precision, const void* L, int64_t N, int64_t ld, double* dinv,
             double* b, double* x, int trans, void* stream);

/* Multi-RHS forward solve L X = S in place on S (N x M, column-major,
 * ld_s rows; N % 64 == 0, M % 64 == 0); storage per `precision`; dinv as
 * for pgp_trsv.  Replaces B.triangular_solve (probit/implicit/VB.py:26-27)
 * and the solve inside predict (probit/approximators.py:177). */
int pgp_trsm(int precision, const void* L, int64_t N, int64_t ld, double* dinv,
             void* S, int64_t M, int64_t ld_s, void* stream);

/* ------------------------------------------------------------------ */
/* Laplace / VB problem entry points                                    */
/* ------------------------------------------------------------------ */

/* Bytes of workspace one problem of N points needs (holds K, the factor,
 * Newton vectors and the device descriptor). */
int64_t pgp_workspace_size(int64_t N, int32_t D, int32_t precision);

/* LaplaceGP.approximate_posterior (probit/approximators.py:204-210):
 * Newton fixed point of a = grad log p(y | K a) from a = 0
 * (probit/approximators.py:265-269, probit/implicit/solvers.py:18-25) and
 * W = -d2 log p(y | K a) (probit/approximators.py:271-277).  Leaves
 * B = I + W^1/2 K W^1/2 factored in the workspace for predict / objective /
 * gradient.  Synchronises the stream once per Newton iteration. */
int pgp_laplace_fit(const pgp_problem_t* problem, const pgp_kernel_t* kernel,
                    const pgp_likelihood_t* lik, const pgp_fit_options_t* opts,
                    pgp_fit_result_t* result, void* workspace,
                    int64_t workspace_bytes, void* stream);

/* pgp_laplace_fit without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops and
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * workspace is a static device array inside the library, one per GPU, held
 * until the process ends (pgp_noalloc_scratch_bytes() bytes: the workspace
 * of N = PGP_NOALLOC_MAX_NP at any precision); larger N gives a negative
 * status.  The call allocates nothing: no hipMalloc / hipFree, no
 * allocating library routine.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other threads) race,
 * and the resident factor a call leaves is overwritten by the next call (no
 * predict / objective from it: use pgp_laplace_fit for those).  Outputs,
 * status and result as pgp_laplace_fit. */
#define PGP_NOALLOC_MAX_NP 1024
int64_t pgp_noalloc_scratch_bytes(void);
int pgp_laplace_fit_noalloc(const pgp_problem_t* problem, const pgp_kernel_t* kernel,
                            const pgp_likelihood_t* lik, const pgp_fit_options_t* opts,
                            pgp_fit_result_t* result, void* stream);

/* Re-form and factor B from a caller-given (weight, precision) pair, so
 * that predict / objective use exactly those values
 * (probit/approximators.py:175 builds K + diag(1/precision)).  With
 * W == NULL the precision is -d2 log p(y | K weight), as in
 * LaplaceGP.precision (probit/approximators.py:271-277); problem->precision
 * and problem->latent (if set) receive W and f = K weight. */
int pgp_laplace_set_state(const pgp_problem_t* problem,
                          const pgp_kernel_t* kernel,
                          const pgp_likelihood_t* lik, int32_t precision,
                          const double* weight, const double* W,
                          pgp_fit_result_t* result, void* workspace,
                          int64_t workspace_bytes, void* stream);

/* Predict workspace bytes for M test points. */
int64_t pgp_predict_workspace_size(int64_t N, int64_t M, int32_t precision);

/* Approximator.predict (probit/approximators.py:154-180) from the resident
 * factor: mean = Kfs^T a, var = Kss - ||L_B^-1 W^1/2 Kfs||^2 (latent).
 * With a factor stored in FP32 (modes fp32, bf16) the variance is corrected
 * in FP64: var = Kss - (2 s^T z - z^T B z), z = (L L^T)^-1 s from the FP32
 * solves and B z from the FP64 Gram matrix, whose error is quadratic in
 * that of z (PGP_PREDICT_REFINE=0 in the environment turns it off).
 * `kernel` must be the resident fit's kernel (the reference's predict takes
 * the parameters it uses): other parameters, or a workspace that holds no
 * predict state (nothing fitted, the VB map's sigma^2 I + K), give a
 * negative status.  Waits for its work; returns PGP_INFO_NOT_PD /
 * PGP_INFO_STALLED (positive) if the factor or the solve is invalid. */
int pgp_predict(const pgp_kernel_t* kernel, const double* Xtest, int64_t M,
                double* mean, double* var, void* workspace,
                int64_t workspace_bytes, void* predict_workspace,
                int64_t predict_workspace_bytes, void* stream);

/* Posterior predictive covariance at M test points from the resident factor
 * (replaces Approximator.predict_covariance, probit/approximators.py:182-197):
 * cov = Kss - Kfs^T (K + W^-1)^-1 Kfs, column-major M x M FP64 with leading
 * dimension ld >= M, computed as Kss - V^T V with V = L_B^-1 W^1/2 Kfs (the
 * same triangular solve as pgp_predict; FP64 MFMA product) and written
 * exactly symmetric.  Workspaces, the kernel check and the status as for
 * pgp_predict. */
int pgp_predict_covariance(const pgp_kernel_t* kernel, const double* Xtest,
                           int64_t M, double* cov, int64_t ld, void* workspace,
                           int64_t workspace_bytes, void* predict_workspace,
                           int64_t predict_workspace_bytes, void* stream);

/* Gradient workspace bytes. */
int64_t pgp_grad_workspace_size(int64_t N, int32_t D);

/* d objective_LA / d(variance, lengthscale[0..D), period[0..D),
 * noise_std, cutpoints[0..J]) at the resident fixed point: the value
 * jax.value_and_grad yields through fixed_point_layer
 * (probit/approximators.py:132-134, probit/implicit/solvers.py:52-64).
 * grad_out is HOST memory of 2 + 2*D + PGP_MAX_CUT doubles laid out as
 * [variance, lengthscale[D], period[D], noise_std, cutpoints[J+1]].
 * `kernel` and `lik` must be the resident Laplace fit's (else a negative
 * status); a non-positive pivot or an expired wait gives NaN gradients and
 * returns PGP_INFO_NOT_PD / PGP_INFO_STALLED. */
int pgp_laplace_grad(const pgp_kernel_t* kernel, const pgp_likelihood_t* lik,
                     double* grad_out, void* workspace, int64_t workspace_bytes,
                     void* grad_workspace, int64_t grad_workspace_bytes,
                     void* stream);

/* VBGP.approximate_posterior (probit/approximators.py:332-339): plain
 * fixed-point iteration of f_VB (probit/implicit/VB.py:4-16) from zero and
 * precision = 1/sigma^2; leaves the factor of I + K / sigma^2 (what predict
 * uses with precision 1/sigma^2) in the workspace.
 * result->objective is objective_VB (probit/implicit/VB.py:19-40), which
 * needs grad_workspace (pgp_grad_workspace_size) for the explicit inverse. */
int pgp_vb_fit(const pgp_problem_t* problem, const pgp_kernel_t* kernel,
               const pgp_likelihood_t* lik, const pgp_fit_options_t* opts,
               pgp_fit_result_t* result, void* workspace,
               int64_t workspace_bytes, void* grad_workspace,
               int64_t grad_workspace_bytes, void* stream);

/* d objective_VB / d(variance, lengthscale[0..D), period[0..D), noise_std,
 * cutpoints[0..J]) at the resident VB fixed point (after pgp_vb_fit, with
 * its workspace and gradient workspace), laid out as pgp_laplace_grad's:
 * the value jax.value_and_grad yields for VBGP through fixed_point_layer
 * (probit/approximators.py:132-134,316-330, probit/implicit/solvers.py:52-64,
 * probit/implicit/VB.py:4-40).  A likelihood curvature below zero somewhere
 * (the probit floor) gives NaN.  The resident state (the factor predict
 * uses) is the same afterwards.  Argument checks and status as for
 * pgp_laplace_grad, against the resident VB fit. */
/* The VB fixed-point map f_VB(w) = (sigma^2 I + K)^-1 (K w + sigma
 * d log p(y | K w) / df) into problem->weight (probit/implicit/VB.py:4-16,
 * VBGP.construct, probit/approximators.py:306-314).  Leaves sigma^2 I + K
 * factored in the workspace (a later predict re-forms its own factor). */
int pgp_vb_map(const pgp_problem_t* problem, const pgp_kernel_t* kernel,
               const pgp_likelihood_t* lik, int32_t precision,
               const double* weight, void* workspace, int64_t workspace_bytes,
               void* stream);

int pgp_vb_grad(const pgp_kernel_t* kernel, const pgp_likelihood_t* lik,
                double* grad_out, void* workspace, int64_t workspace_bytes,
                void* grad_workspace, int64_t grad_workspace_bytes,
                void* stream);

/* ------------------------------------------------------------------ */
/* Batched serving entry points                                         */
/* ------------------------------------------------------------------ */

/* One batch member: a problem, its kernel and likelihood, its workspace
 * and (for predict) its test inputs and outputs. */
typedef struct pgp_batch_item {
  pgp_problem_t problem;
  pgp_kernel_t kernel;
  pgp_likelihood_t lik;
  void* workspace;
  int64_t workspace_bytes;
  const double* Xtest; /* M x D row-major, may be NULL (no predict) */
  int64_t M;
  double* mean;
  double* var;
  void* predict_workspace;
  int64_t predict_workspace_bytes;
  pgp_fit_result_t result; /* output */
} pgp_batch_item_t;

/* Continuous-batching engine (library-owned host object; its device tables
 * live in a caller-owned block of pgp_batcher_table_size(capacity) bytes).
 * submit() queues an item (its memory must stay valid until its tag is
 * returned); each step() admits queued items up to `capacity`, runs ONE
 * Newton iteration for every item in flight with one launch per blocked
 * step shared by all of them, then finalises and predicts the items that
 * converged and returns their tags.  Items join and leave between steps.
 * An item's results do not depend on which other items share its steps.
 * The steps run on streams the batcher creates (default priority); each
 * step first waits for what the caller had enqueued on `stream` when it was
 * started (the items' inputs), and step() returns (finish() reports) only
 * after the step's work is complete, so the results are ready for any
 * later work.  This is the batched form of LaplaceGP.approximate_posterior + predict
 * (probit/approximators.py:204-210, 154-180). */
typedef struct pgp_batcher pgp_batcher_t;
int64_t pgp_batcher_table_size(int32_t capacity);
int pgp_batcher_create(int32_t capacity, const pgp_fit_options_t* opts,
                       void* table, int64_t table_bytes, void* stream,
                       pgp_batcher_t** out);
int pgp_batcher_submit(pgp_batcher_t* batcher, pgp_batch_item_t* item,
                       int64_t tag);
int pgp_batcher_step(pgp_batcher_t* batcher, int64_t* done_tags,
                     int32_t max_done, int32_t* n_done);
/* Asynchronous form of pgp_batcher_step, so that the host can prepare and
 * submit requests while the GPU works.  Items are split by padded size into
 * four lanes (bounds 1024, 1536 and 2560, or PGP_LANE_NP in the
 * environment; the lanes past the second take items only while at least 512
 * are in flight, PGP_LANE_SPLIT_MIN), each with its own streams and step
 * cadence, so that the small items run several steps while the large ones
 * run one.  launch() admits
 * queued items and, on every lane without a step in flight, enqueues one
 * step (the finalisation and predict of the lane's items that converged in
 * its previous step, one Newton iteration for its other items), returning
 * at once; finish() waits until at least one lane's step is done and returns
 * the tags of the items it completed; launched() counts the lanes with a
 * step in flight; ready() counts the lanes whose step is done (finish()
 * would return at once) without waiting.  An item is reported one step
 * after it converges; its
 * results are the same as with pgp_batcher_step.  Do not call
 * pgp_batcher_step while a lane has a step in flight. */
int pgp_batcher_launch(pgp_batcher_t* batcher);
int pgp_batcher_finish(pgp_batcher_t* batcher, int64_t* done_tags,
                       int32_t max_done, int32_t* n_done);
int32_t pgp_batcher_launched(pgp_batcher_t* batcher);
int32_t pgp_batcher_ready(pgp_batcher_t* batcher);
int32_t pgp_batcher_in_flight(pgp_batcher_t* batcher);
/* Host-side counters since creation, for measurement: seconds spent in
 * pgp_batcher_launch (admission and enqueueing of lane steps), seconds
 * pgp_batcher_finish waited for a lane's step to complete, seconds it spent
 * reporting completed steps, and the number of lane steps launched.  Any
 * output may be NULL.  No counterpart in the reference (measurement only). */
int pgp_batcher_stats(pgp_batcher_t* batcher, double* host_enqueue_s, double* host_wait_s,
                      double* host_report_s, int64_t* lane_steps);
int pgp_batcher_destroy(pgp_batcher_t* batcher);

#ifdef __cplusplus
}
#endif

#endif /* PROBIT_AMD_H */
