// Host-side drivers and the extern "C" ABI of libprobit_amd.
//
// Control flow mirrors the reference call stacks:
//   pgp_laplace_fit   <- LaplaceGP.approximate_posterior (probit/approximators.py:204-210)
//                        Newton loop of newton_solver / fwd_solver (probit/implicit/solvers.py:7-25)
//   pgp_predict       <- Approximator.predict (probit/approximators.py:154-180)
//   pgp_laplace_grad  <- value_and_grad(objective) (probit/approximators.py:132-134, 248-263)
//   pgp_vb_fit        <- VBGP.approximate_posterior + objective_VB (probit/approximators.py:306-339)
// The stopping rule is jaxopt's: iterate until ||z_{k+1} - z_k||_2 <= tol or
// maxiter, returning z_{k+1}.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <list>
#include <atomic>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pgp_internal.hpp"

namespace pgp {

// grad.hip
void launch_sym_from_inv(const Batch& b, int prec, int use_scale, hipStream_t s);
void launch_fill_rhs(const Batch& b, int prec, int mode, hipStream_t s);
void launch_sigma_diag(const Batch& b, int prec, hipStream_t s);
void launch_grad_trace(const Batch& b, double* partials, double* out, hipStream_t s);
void launch_vb_trace(const Batch& b, double* out, hipStream_t s);
void launch_vb_sigma_terms(const Batch& b, double* out, hipStream_t s);
int grad_npar();
int grad_maxd();

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_OK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) return fail(-100, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

static int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(-101, "kernel launch: %s", hipGetErrorString(e));
  return 0;
}

static inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }
static inline int64_t align256(int64_t x) { return round_up(x, 256); }
static inline int elem_size(int prec) { return prec == PGP_PREC_FP64 ? 8 : 4; }

// ---------------------------------------------------------------------------
// workspace layouts
// ---------------------------------------------------------------------------
constexpr int64_t OFF_PROB = 0;
constexpr int64_t OFF_ARR = 2048;   // pointer array (self) for single-problem launches
constexpr int64_t OFF_INFO = 2112;  // int32[4]
constexpr int64_t OFF_RED = 2176;   // double[NRED]
constexpr int64_t OFF_GATHER = 2432;  // double[8] scratch for reductions
constexpr int64_t OFF_PROGRESS = 2496;  // uint64 progress word of the wavefront solves
constexpr int64_t OFF_TICKET = 2504;    // uint32 workgroup-order counter (single-problem launches)
constexpr int64_t OFF_PROGRESS_CT = 2560;  // uint64[4] progress words of the wavefront TRSM
constexpr int64_t HEADER = 4096;

struct WsLayout {
  int64_t Np, K, L, logdiag, diagval, dinv, vec, symv, csync, panel16, total;
};
// Columns of the BF16 panel copy: a super-step is at most 16 block columns.
static inline int64_t panel16_pitch(int64_t Np) { return std::min<int64_t>(Np, 16 * TILE); }

static WsLayout ws_layout(int64_t N, int prec) {
  WsLayout w;
  w.Np = round_up(N, TILE);
  int64_t off = HEADER;
  w.K = off;
  off += align256(w.Np * w.Np * 8);
  w.L = off;
  off += align256(w.Np * w.Np * elem_size(prec));
  w.logdiag = off;
  off += align256(w.Np * 8);
  w.diagval = off;
  off += align256(w.Np * 8);
  w.dinv = off;
  off += align256(w.Np * TILE * 8);
  w.vec = off;
  off += align256((int64_t)NVEC * w.Np * 8);
  w.symv = off;
  off += align256(w.Np * (w.Np / TILE) * 8);
  w.csync = off;
  off += align256(csync_words(w.Np) * 4);
  w.panel16 = 0;
  if (prec == PGP_PREC_BF16) {
    w.panel16 = off;
    off += align256(w.Np * panel16_pitch(w.Np) * 2);
  }
  w.total = off;
  return w;
}

struct PredLayout {
  int64_t Mp, S, mupart, qpart, total;
};
static PredLayout pred_layout(int64_t Np, int64_t M, int prec) {
  PredLayout p;
  p.Mp = round_up(M, TILE);
  int64_t off = 0;
  p.S = off;
  off += align256(Np * p.Mp * elem_size(prec));
  p.mupart = off;
  off += align256((Np / TILE) * p.Mp * 8);
  p.qpart = off;
  off += align256((Np / TILE) * p.Mp * 8);
  p.total = off;
  return p;
}

// Factors stored in FP32 (modes fp32, bf16) get the FP64 correction of the
// predictive variance (k_pred_refine) unless PGP_PREDICT_REFINE=0.
static bool predict_refine(int prec) {
  static const bool on = [] {
    const char* e = getenv("PGP_PREDICT_REFINE");
    return !(e && e[0] == '0');
  }();
  return on && prec != PGP_PREC_FP64;
}

// The predict of the members of `b` (descriptors with S, mupart, qpart and
// the outputs set): S = W^1/2 Kfs with the partial means, S <- L^-1 S, and
// either var = Kss - ||L^-1 s||^2 or, with the correction, S <- L^-T S and
// the FP64 quadratic forms.
static void predict_launches(const Batch& b, int prec, int maxM, hipStream_t s) {
  const int maxMp = (int)round_up(maxM, TILE);
  launch_cross_prep(b, prec, maxM, s);
  launch_trsm(b, prec, maxMp, 0, s);
  if (predict_refine(prec)) {
    launch_trsm_t(b, prec, maxMp, s);
    launch_pred_refine(b, prec, maxM, s);
  }
  launch_predict_final(b, prec, maxM, s);
}

struct GradLayout {
  int64_t C, G2, partials, out, total;
};
static GradLayout grad_layout(int64_t Np, int prec) {
  GradLayout g;
  const int64_t T = Np / TILE;
  int64_t off = 0;
  g.C = off;
  off += align256(Np * Np * 8);
  g.G2 = off;
  off += align256(Np * Np * elem_size(prec));
  g.partials = off;
  off += align256(T * (T + 1) / 2 * grad_npar() * 8);
  g.out = off;
  off += align256(256 * 8);
  g.total = off;
  return g;
}

// ---------------------------------------------------------------------------
// device-side descriptor and pointer-array writes (kernel arguments carry the
// host values, so no host buffer has to outlive the call)
// ---------------------------------------------------------------------------
__global__ void k_write_prob(Prob* dst, Prob val, int reset) {
  const uint64_t* src = reinterpret_cast<const uint64_t*>(&val);
  uint64_t* d = reinterpret_cast<uint64_t*>(dst);
  const int nw = (int)(sizeof(Prob) / 8);
  for (int i = threadIdx.x; i < nw; i += blockDim.x) d[i] = src[i];
  if (reset) {
    if (threadIdx.x < 4) val.info[threadIdx.x] = 0;
    if (threadIdx.x < NRED) val.red[threadIdx.x] = 0.0;
    // progress word and ticket counter of the header
    if (threadIdx.x < 2) reinterpret_cast<uint64_t*>((char*)dst + OFF_PROGRESS)[threadIdx.x] = 0;
    if (threadIdx.x < 4) reinterpret_cast<uint64_t*>((char*)dst + OFF_PROGRESS_CT)[threadIdx.x] = 0;
    if (val.csync)
      for (int64_t w = threadIdx.x; w < csync_words(val.Np); w += blockDim.x) val.csync[w] = 0;
  }
}

constexpr int PTRPACK = 448;
struct PtrPack {
  Prob* p[PTRPACK];
};
__global__ void k_write_ptrs(Prob** dst, int n, PtrPack pk) {
  const int i = threadIdx.x + blockIdx.x * blockDim.x;
  if (i < n) dst[i] = pk.p[i];
}

static_assert(sizeof(Prob) % 8 == 0, "Prob must be 8-byte granular");
static_assert(sizeof(Prob) <= 2048, "Prob must fit the workspace header");

// Many descriptors at once: `src` (device) holds n staged descriptors,
// dst[i] is where descriptor i goes.
__global__ void k_scatter_probs(Prob* const* dst, const Prob* src, int n, int reset) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const uint64_t* sv = reinterpret_cast<const uint64_t*>(src + i);
  uint64_t* d = reinterpret_cast<uint64_t*>(dst[i]);
  const int nw = (int)(sizeof(Prob) / 8);
  for (int w = threadIdx.x; w < nw; w += blockDim.x) d[w] = sv[w];
  if (reset) {
    const Prob& v = src[i];
    if (threadIdx.x < 4) v.info[threadIdx.x] = 0;
    if (threadIdx.x < NRED) v.red[threadIdx.x] = 0.0;
    if (threadIdx.x == 0) v.progress[0] = 0;
    if (threadIdx.x < 4) v.progress_ct[threadIdx.x] = 0;
    if (v.csync)
      for (int64_t w = threadIdx.x; w < csync_words(v.Np); w += blockDim.x) v.csync[w] = 0;
  }
}

static void write_prob(char* ws, const Prob& p, int reset, hipStream_t s) {
  hipLaunchKernelGGL(k_write_prob, dim3(1), dim3(256), 0, s, (Prob*)(ws + OFF_PROB), p, reset);
}

static void write_ptrs(Prob** dst, const std::vector<Prob*>& v, hipStream_t s) {
  for (size_t b = 0; b < v.size(); b += PTRPACK) {
    PtrPack pk;
    const int n = (int)std::min<size_t>(PTRPACK, v.size() - b);
    for (int i = 0; i < n; ++i) pk.p[i] = v[b + i];
    hipLaunchKernelGGL(k_write_ptrs, dim3((n + 255) / 256), dim3(256), 0, s, dst + b, n, pk);
  }
}

// ---------------------------------------------------------------------------
// descriptor construction from the ABI structs
// ---------------------------------------------------------------------------
static int validate(const pgp_problem_t* pr, const pgp_kernel_t* k, const pgp_likelihood_t* l) {
  if (!pr) return fail(-1, "problem is NULL");
  if (pr->N <= 0) return fail(-1, "problem N must be positive (got %lld)", (long long)pr->N);
  if (pr->D <= 0 || pr->D > PGP_MAX_DIM)
    return fail(-1, "problem D must be in [1, %d] (got %d)", PGP_MAX_DIM, pr->D);
  if (!pr->X || !pr->y) return fail(-1, "problem X / y must be device pointers");
  if (!k) return fail(-2, "kernel is NULL");
  if (k->kind != PGP_KERNEL_EQ && k->kind != PGP_KERNEL_MATERN12)
    return fail(-2, "unknown kernel kind %d", k->kind);
  if (k->dim != pr->D) return fail(-2, "kernel dim %d != problem D %d", k->dim, pr->D);
  if (!(k->variance > 0.0)) return fail(-2, "kernel variance must be > 0");
  for (int d = 0; d < k->dim; ++d) {
    if (!(k->lengthscale[d] > 0.0)) return fail(-2, "lengthscale[%d] must be > 0", d);
    if (k->period[d] < 0.0) return fail(-2, "period[%d] must be >= 0", d);
  }
  if (l) {
    if (l->kind < PGP_LIK_GAUSSIAN || l->kind > PGP_LIK_PROBIT_CLOSED)
      return fail(-3, "unknown likelihood kind %d", l->kind);
    if (!(l->noise_std > 0.0)) return fail(-3, "noise_std must be > 0");
    if (l->kind != PGP_LIK_GAUSSIAN) {
      if (l->J < 2 || l->J + 1 > PGP_MAX_CUT)
        return fail(-3, "J must be in [2, %d] (got %d)", PGP_MAX_CUT - 1, l->J);
      if (!(isinf(l->cutpoints[0]) && l->cutpoints[0] < 0) ||
          !(isinf(l->cutpoints[l->J]) && l->cutpoints[l->J] > 0))
        return fail(-3, "cutpoints must start at -inf and end at +inf");
      for (int j = 0; j < l->J; ++j)
        if (!(l->cutpoints[j] <= l->cutpoints[j + 1]))
          return fail(-3, "cutpoints must be ascending");
    }
  }
  return 0;
}

static void fill_kernel(Prob& p, const pgp_kernel_t* k) {
  p.kern_kind = k->kind;
  p.variance = k->variance;
  p.periodic = 0;
  for (int d = 0; d < PGP_MAX_DIM; ++d) {
    p.ls[d] = 1.0;
    p.inv_ls2[d] = 0.0;
    p.pi_over_per[d] = 0.0;
    p.per[d] = 0.0;
  }
  for (int d = 0; d < k->dim; ++d) {
    p.ls[d] = k->lengthscale[d];
    p.inv_ls2[d] = 1.0 / (k->lengthscale[d] * k->lengthscale[d]);
    if (k->period[d] > 0.0) {
      p.periodic = 1;
      p.per[d] = k->period[d];
      p.pi_over_per[d] = M_PI / k->period[d];
    }
  }
}

static void fill_lik(Prob& p, const pgp_likelihood_t* l) {
  p.lik_kind = l ? l->kind : PGP_LIK_GAUSSIAN;
  p.J = l ? l->J : 0;
  p.sp_bounds = l ? l->single_precision_bounds : 1;
  p.noise_std = l ? l->noise_std : 1.0;
  for (int c = 0; c < PGP_MAX_CUT; ++c) p.cut[c] = l ? l->cutpoints[c] : 0.0;
}

static Prob make_prob(const pgp_problem_t* pr, const pgp_kernel_t* k, const pgp_likelihood_t* l,
                      int prec, char* ws) {
  Prob p;
  memset(&p, 0, sizeof p);
  const WsLayout L = ws_layout(pr->N, prec);
  p.N = (int32_t)pr->N;
  p.Np = (int32_t)L.Np;
  p.D = pr->D;
  p.prec = prec;
  fill_kernel(p, k);
  fill_lik(p, l);
  p.X = pr->X;
  p.y = pr->y;
  p.K = (double*)(ws + L.K);
  p.L = ws + L.L;
  p.logdiag = (double*)(ws + L.logdiag);
  p.diagval = (double*)(ws + L.diagval);
  p.dinv = (double*)(ws + L.dinv);
  p.vec = (double*)(ws + L.vec);
  p.symv = (double*)(ws + L.symv);
  p.csync = (uint32_t*)(ws + L.csync);
  p.panel16 = L.panel16 ? (unsigned short*)(ws + L.panel16) : nullptr;
  p.panel16_pitch = (int32_t)panel16_pitch(L.Np);
  p.red = (double*)(ws + OFF_RED);
  p.info = (int32_t*)(ws + OFF_INFO);
  p.progress = (uint64_t*)(ws + OFF_PROGRESS);
  p.progress_ct = (uint64_t*)(ws + OFF_PROGRESS_CT);
  p.out_weight = pr->weight;
  p.out_prec = pr->precision;
  p.out_latent = pr->latent;
  return p;
}

// Read the descriptor's host-visible fields back from a workspace.
static int read_prob(const char* ws, Prob& p, hipStream_t s) {
  HIP_OK(hipMemcpyAsync(&p, ws + OFF_PROB, sizeof(Prob), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  return 0;
}

// Entry points that act on the resident factor (predict, covariance,
// gradients) compute from the resident descriptor, so the kernel (and, for
// the gradients, likelihood) a caller passes must be the resident fit's:
// the reference's predict(X_test, parameters, ...) and value_and_grad()(parameters)
// use the parameters they are given (probit/approximators.py:171-174, :132-134).
// `states` is the set of ResidentState values (bit per value) the call accepts.
static int check_resident(const Prob& p, unsigned states, const char* who,
                          const pgp_kernel_t* k, int karg, const pgp_likelihood_t* l,
                          int larg) {
  if (!((states >> p.state) & 1u)) {
    static const char* names[] = {"nothing fitted", "a Laplace state", "a VB fit",
                                  "the VB map's sigma^2 I + K"};
    return fail(-99, "%s: the workspace holds %s", who,
                (p.state >= 0 && p.state <= RS_VBMAP) ? names[p.state] : "an unknown state");
  }
  if (!k) return fail(karg, "%s: kernel is NULL", who);
  if (k->dim != p.D) return fail(karg, "%s: kernel dim %d != the resident fit's D %d", who, k->dim, p.D);
  Prob q;
  memset(&q, 0, sizeof q);
  fill_kernel(q, k);
  bool same = q.kern_kind == p.kern_kind && q.variance == p.variance;
  for (int d = 0; d < p.D; ++d) same = same && q.ls[d] == p.ls[d] && q.per[d] == p.per[d];
  if (!same)
    return fail(karg, "%s: kernel parameters differ from the resident fit's (fit again, or "
                      "re-form the state with pgp_laplace_set_state, first)", who);
  if (larg) {
    if (!l) return fail(larg, "%s: likelihood is NULL", who);
    fill_lik(q, l);
    same = q.lik_kind == p.lik_kind && q.noise_std == p.noise_std;
    if (same && q.lik_kind != PGP_LIK_GAUSSIAN) {
      same = q.J == p.J && (q.lik_kind != PGP_LIK_PROBIT_CLOSED || q.sp_bounds == p.sp_bounds);
      for (int c = 0; same && c <= q.J; ++c) same = q.cut[c] == p.cut[c];
    }
    if (!same)
      return fail(larg, "%s: likelihood parameters differ from the resident fit's (fit again "
                        "first)", who);
  }
  return 0;
}

constexpr unsigned RS_PREDICT = (1u << RS_LAPLACE) | (1u << RS_VB);

// Status bits of the factor and of the solves since the state was formed
// (a non-positive pivot, an expired bounded wait): after the call's work.
static int read_info_bits(const char* ws, hipStream_t s, int* bits) {
  int32_t info = 0;
  HIP_OK(hipMemcpyAsync(&info, ws + OFF_INFO, sizeof info, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  *bits = info & (PGP_INFO_NOT_PD | PGP_INFO_STALLED);
  return 0;
}

static Batch single_batch(char* ws, int Np) {
  Batch b;
  b.dprobs = (Prob* const*)(ws + OFF_ARR);
  b.count = 1;
  b.maxNp = Np;
  b.ticket = (uint32_t*)(ws + OFF_TICKET);
  return b;
}

static void set_self(char* ws, hipStream_t s) {
  std::vector<Prob*> v(1, (Prob*)(ws + OFF_PROB));
  write_ptrs((Prob**)(ws + OFF_ARR), v, s);
}

// ---------------------------------------------------------------------------
// Newton pieces (batched)
// ---------------------------------------------------------------------------
static void factor(const Batch& b, int prec, int vb, hipStream_t s, int rs = -1) {
  launch_form_B(b, (prec << 8) | vb, s, rs);
  launch_potrf(b, prec, s);
  launch_dinv(b, prec, s);
}

static void solve_B(const Batch& b, int prec, hipStream_t s) {
  launch_trsv(b, prec, 0, V_T, V_X3, s);
  launch_trsv(b, prec, 1, V_X3, V_U, s);
}

// One Newton iteration z <- z + (I + W K)^-1 (grad - z) for every member of
// `all` (`later`: those past their first iteration); members of `fac`
// re-form and re-factor B first (those of `fac32`
// with FP32 updates: BF16 members escalated after a loss of definiteness;
// both store the factor in FP32), and `nofac` lists the others.  The pass
// that forms B also takes the partial sums of K r, so only the members of
// `nofac` read K once more for it.  ||delta||^2 goes to gather[b].
static void newton_iter(const Batch& all, const Batch* later, const Batch* fac,
                        const Batch* fac32, const Batch* nofac, int prec, int refine,
                        double* gather, hipStream_t s) {
  // f = K z; at a member's first iteration z = 0 and f was zeroed with it
  if (later && later->count > 0) launch_gemv_sym(*later, V_Z, V_F, 0, s);
  launch_lik_newton(all, s);
  // the members that re-factor B solve with the new W^1/2; the others keep
  // the one their factor was formed with (a chord step)
  if (fac && fac->count > 0) launch_vec_op(*fac, OP_SNAP, s);
  if (fac32 && fac32->count > 0) launch_vec_op(*fac32, OP_SNAP, s);
  if (fac && fac->count > 0) factor(*fac, prec, 0, s, V_R);
  if (fac32 && fac32->count > 0) factor(*fac32, PGP_PREC_FP32, 0, s, V_R);
  if (nofac && nofac->count > 0) launch_symv_tiles(*nofac, V_R, s);
  launch_symv_reduce(all, V_R, V_T, 1, s);
  solve_B(all, prec, s);
  launch_vec_op(all, OP_STEP, s);
  for (int r = 0; r < refine; ++r) {
    launch_gemv_sym(all, V_DL, V_Q, 2, s);
    launch_gemv_sym(all, V_Q, V_T, 1, s);
    solve_B(all, prec, s);
    launch_vec_op(all, OP_IRCORR, s);
  }
  launch_vec_op(all, OP_ZUPD, s);
  launch_reduce(all, RED_STEP, gather, s);
}

// Final state at the returned weight: f, W, sum log p, and B factored at W.
// In the BF16 mode the Newton iterations factor with BF16-operand updates
// (the FP64 residual keeps their fixed point the FP64 one), but the final
// factor -- the one predict, the log-determinant and the gradient use -- is
// formed with FP32 updates for every member (final_fp32): a BF16-update
// factor's error (~1e-3 relative) would pass into the predictive variance
// and the objective.  The single-problem fit and the batcher apply the same
// rule, so their results stay bitwise equal.
static void laplace_finalize(const Batch& all, const Batch* fac, const Batch* fac32, int prec,
                             double* gather, hipStream_t s) {
  launch_gemv_sym(all, V_Z, V_F, 0, s);
  launch_lik_final(all, 0, s);
  if (fac && fac->count > 0) factor(*fac, prec, 0, s);
  if (fac32 && fac32->count > 0) factor(*fac32, PGP_PREC_FP32, 0, s);
  launch_reduce(all, RED_FINAL, gather, s);
  launch_vec_op(all, OP_OUT, s);
}

// Status bits of a fit from its device status word, and a NaN objective
// where the bits make the numbers invalid (as JAX would give NaN).
static void apply_info(pgp_fit_result_t* r, int info) {
  if (info & PGP_INFO_NOT_PD) r->status |= PGP_INFO_NOT_PD;
  if (info & PGP_INFO_STALLED) r->status |= PGP_INFO_STALLED;
  if (r->status & (PGP_INFO_NEGATIVE_W | PGP_INFO_NOT_PD | PGP_INFO_STALLED)) r->objective = NAN;
}

static void fill_result_final(pgp_fit_result_t* r, const double* g4, int info) {
  r->sum_log_lik = g4[0];
  r->half_f_dot_a = 0.5 * g4[1];
  r->log_det = 2.0 * g4[2];
  if (g4[3] > 0) r->status |= PGP_INFO_NEGATIVE_W;
  r->objective = -g4[0] + 0.5 * g4[1] + g4[2];
  apply_info(r, info);
}

// Clear only the non-positive-pivot bit of each member's status word (a
// stall flagged in the same iteration must survive into the final status).
__global__ void k_clear_not_pd(Prob* const* arr, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAnd(arr[i]->info, ~(int32_t)PGP_INFO_NOT_PD);
}
static void launch_clear_not_pd(const Batch& b, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_not_pd, dim3((b.count + 255) / 256), dim3(256), 0, s, b.dprobs,
                     b.count);
}

// FP64 refinement steps per Newton solve by default.  The Newton residual is
// always FP64, so an FP32 factor only slows the (inexact) Newton contraction:
// no refinement costs ~10% more iterations but half the solves per
// iteration.  BF16 trailing updates need the refinement.
static int default_refine(int prec) { return prec == PGP_PREC_BF16 ? 2 : 0; }

// Inexact-Newton stall test of the BF16 mode: from the third iteration on, a
// step norm above half the previous one.
static bool stalled(int iteration, double step, double prev) {
  return iteration >= 3 && step > 0.5 * prev;
}

// Newton iterations of the FP32 mode re-use the factor of B once the step
// norm has fallen to CHORD_STEP (a chord step: B and its W^1/2 from the last
// factorisation, the residual at the current weight in FP64, so the fixed
// point is the same); a chord step that contracts by less than half, or a
// larger step, re-factors at the next iteration.  The FP64 and BF16 modes
// re-factor at every iteration of a non-Gaussian likelihood (the
// reference's Newton, probit/implicit/solvers.py:18-25), a Gaussian one only
// at the first (W is constant).
static double env_double(const char* name, double dflt) {
  const char* e = getenv(name);
  return e ? atof(e) : dflt;
}
// (PGP_CHORD_STEP = 0 turns the chord steps off)
static const double CHORD_STEP = env_double("PGP_CHORD_STEP", 0.1);
// A run of chord steps converges linearly, so it stops at a tenth of the
// tolerance (a Newton step's error is quadratic in its step norm).
static const double CHORD_TOL = env_double("PGP_CHORD_TOL", 0.1);
struct NewtonTrack {
  double prev = INFINITY, prev2 = INFINITY;  // the last two step norms
  bool chord = false;                        // the last iteration re-used the factor
};
static bool refactor_now(int prec, bool gauss, int it, bool force, const NewtonTrack& t) {
  if (it == 0 || force) return true;
  if (gauss) return false;
  if (prec != PGP_PREC_FP32) return true;
  if (!(t.prev <= CHORD_STEP)) return true;
  return t.chord && t.prev > 0.5 * t.prev2;
}
// A chord step: an FP32-mode iteration of a non-Gaussian likelihood that
// re-used the factor (a Gaussian likelihood's W never changes).
static bool chord_step(int prec, bool gauss, bool refactored) {
  return !refactored && !gauss && prec == PGP_PREC_FP32;
}
static void track_step(NewtonTrack& t, double step, bool chord) {
  t.prev2 = t.prev;
  t.prev = step;
  t.chord = chord;
}
static bool newton_done(double step, double tol, const NewtonTrack& t) {
  return step <= (t.chord ? CHORD_TOL * tol : tol);
}

// ---------------------------------------------------------------------------
// opt-in launch timer (HIP events on the launching stream)
// ---------------------------------------------------------------------------
struct TimerRec {
  hipEvent_t a, b;
  double flops, bytes;
  int cls;
};
struct TimerState {
  unsigned mask = 0;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<TimerRec> recs;
  hipEvent_t open[TC_COUNT] = {};
};
static TimerState& tstate() {
  static TimerState t;
  return t;
}
bool timer_on(int cls) { return (tstate().mask >> cls) & 1u; }
static hipEvent_t timer_event() {
  TimerState& t = tstate();
  if (t.used == t.pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    t.pool.push_back(e);
  }
  return t.pool[t.used++];
}
void timer_begin(int cls, hipStream_t s) {
  hipEvent_t e = timer_event();
  if (e) (void)hipEventRecord(e, s);
  tstate().open[cls] = e;
}
void timer_end(int cls, hipStream_t s, double flops, double bytes) {
  hipEvent_t e = timer_event();
  if (e) (void)hipEventRecord(e, s);
  TimerState& t = tstate();
  if (t.open[cls] && e) t.recs.push_back({t.open[cls], e, flops, bytes, cls});
  t.open[cls] = nullptr;
}

}  // namespace pgp

using namespace pgp;

extern "C" {

const char* pgp_last_error(void) { return g_err.c_str(); }

int pgp_timer_enable(uint32_t mask) {
  tstate().mask = mask;
  return 0;
}

int pgp_timer_reset(void) {
  TimerState& t = tstate();
  t.recs.clear();
  t.used = 0;
  return 0;
}

int pgp_timer_read(int32_t cls, int64_t* launches, double* total_ms, double* algo_flops,
                   double* algo_bytes) {
  if (cls < 0 || cls >= TC_COUNT) return fail(-1, "unknown timer class %d", cls);
  TimerState& t = tstate();
  int64_t n = 0;
  double ms = 0.0, fl = 0.0, by = 0.0;
  for (const TimerRec& r : t.recs) {
    if (r.cls != cls) continue;
    HIP_OK(hipEventSynchronize(r.b));
    float dt = 0.f;
    HIP_OK(hipEventElapsedTime(&dt, r.a, r.b));
    ++n;
    ms += dt;
    fl += r.flops;
    by += r.bytes;
  }
  if (launches) *launches = n;
  if (total_ms) *total_ms = ms;
  if (algo_flops) *algo_flops = fl;
  if (algo_bytes) *algo_bytes = by;
  return 0;
}
int pgp_version(void) { return 2; }

int pgp_tune_get(int32_t precision, int32_t size_class, pgp_tune_entry_t* out) {
  if (precision < 0 || precision > 2) return fail(-1, "unknown precision %d", precision);
  if (size_class < 0 || size_class >= PGP_TUNE_CLASSES)
    return fail(-2, "size class must be in [0, %d)", PGP_TUNE_CLASSES);
  if (!out) return fail(-3, "out is NULL");
  *out = tune_get(precision, size_class);
  return 0;
}

int pgp_tune_set(int32_t precision, int32_t size_class, const pgp_tune_entry_t* entry) {
  if (precision < 0 || precision > 2) return fail(-1, "unknown precision %d", precision);
  if (size_class < 0 || size_class >= PGP_TUNE_CLASSES)
    return fail(-2, "size class must be in [0, %d)", PGP_TUNE_CLASSES);
  if (!entry) return fail(-3, "entry is NULL");
  if (entry->superstep < 1 || entry->superstep > 16)
    return fail(-3, "superstep must be in [1, 16] (got %d)", entry->superstep);
  if (entry->column_max < 0 || entry->fuse_max < 0)
    return fail(-3, "column_max / fuse_max must be >= 0");
  if (entry->column_merge < 0 || entry->column_merge > 2)
    return fail(-3, "column_merge must be 0, 1 or 2 (got %d)", entry->column_merge);
  tune_set(precision, size_class, *entry);
  return 0;
}

int pgp_fit_result_apply_info(pgp_fit_result_t* result, int32_t info_word) {
  if (!result) return fail(-1, "result is NULL");
  apply_info(result, info_word);
  return 0;
}

int pgp_gram(const pgp_kernel_t* kernel, int dtype, const double* X, int64_t N, const double* Z,
             int64_t M, void* out, int64_t ld, void* stream) {
  if (!kernel) return fail(-1, "kernel is NULL");
  if (dtype != 0 && dtype != 1) return fail(-2, "dtype must be 0 (FP64) or 1 (FP32)");
  if (!X || N <= 0) return fail(-3, "X / N invalid");
  if (!Z || M <= 0) return fail(-5, "Z / M invalid");
  if (!out) return fail(-7, "out is NULL");
  if (ld < N) return fail(-8, "ld (%lld) < N (%lld)", (long long)ld, (long long)N);
  if (kernel->dim <= 0 || kernel->dim > PGP_MAX_DIM) return fail(-1, "kernel dim out of range");
  launch_gram_raw(kernel, dtype, X, N, Z, M, out, ld, (hipStream_t)stream);
  return check_launch();
}

int pgp_likelihood_eval(const pgp_likelihood_t* lik, const double* f, const double* y, int64_t N,
                        double* ll, double* d1, double* d2, void* stream) {
  if (!lik) return fail(-1, "likelihood is NULL");
  if (!f || !y) return fail(-2, "f / y are NULL");
  if (N < 0) return fail(-4, "N < 0");
  launch_lik_raw(lik, f, y, N, ll, d1, d2, (hipStream_t)stream);
  return check_launch();
}

int pgp_probit_predictive(const pgp_likelihood_t* lik, const double* mean, const double* var,
                          int64_t M, double* out, void* stream) {
  if (!lik) return fail(-1, "likelihood is NULL");
  if (lik->kind == PGP_LIK_GAUSSIAN) return fail(-1, "needs a probit likelihood");
  if (lik->J < 2 || lik->J + 1 > PGP_MAX_CUT) return fail(-1, "J out of range");
  if (!mean || !var || !out) return fail(-2, "NULL pointer");
  if (M < 0) return fail(-4, "M < 0");
  launch_probit_predictive(lik, mean, var, M, out, (hipStream_t)stream);
  return check_launch();
}

int pgp_ordinal_metrics(const double* P, int32_t J, const double* y_test, int64_t M,
                        int32_t* y_pred, double* predictive_likelihood, double* metrics,
                        double* workspace, void* stream) {
  if (!P) return fail(-1, "P is NULL");
  if (J < 2 || J + 1 > PGP_MAX_CUT) return fail(-2, "J must be in [2, %d]", PGP_MAX_CUT - 1);
  if (!y_test) return fail(-3, "y_test is NULL");
  if (M <= 0 || M > (int64_t)INT32_MAX) return fail(-4, "M out of range");
  if (!y_pred || !metrics) return fail(-5, "y_pred / metrics are NULL");
  if (!workspace) return fail(-8, "workspace (3 M doubles) is NULL");
  launch_ordinal_metrics(P, J, y_test, M, y_pred, predictive_likelihood, metrics, workspace,
                         (hipStream_t)stream);
  return check_launch();
}

int pgp_potrf(int precision, void* A, int64_t N, int64_t ld, double* dinv, double* logdiag,
              int32_t* info, void* stream) {
  if (precision < 0 || precision > 2) return fail(-1, "unknown precision %d", precision);
  if (!A) return fail(-2, "A is NULL");
  if (N <= 0 || N % TILE) return fail(-3, "N must be a positive multiple of 64");
  if (ld < N) return fail(-4, "ld < N");
  if (!dinv) return fail(-5, "dinv must be a device buffer of N * 64 doubles");
  if (!info) return fail(-7, "info must be a device int32");
  launch_potrf_raw(precision, A, N, ld, dinv, logdiag, info, (hipStream_t)stream);
  return check_launch();
}

int64_t pgp_prior_sample_workspace_size(int64_t N) {
  if (N <= 0) return -1;
  return prior_ws_bytes((N + TILE - 1) / TILE * TILE);
}

int pgp_prior_sample(const pgp_kernel_t* kernel, const double* X, int64_t N, double jitter,
                     const double* z, int64_t S, int64_t ldz, double* f, int64_t ldf,
                     void* workspace, int64_t workspace_bytes, void* stream) {
  if (!kernel) return fail(-1, "kernel is NULL");
  if (kernel->dim <= 0 || kernel->dim > PGP_MAX_DIM) return fail(-1, "kernel dim out of range");
  if (!X || N <= 0) return fail(-2, "X / N invalid");
  if (!(jitter >= 0.0)) return fail(-4, "jitter must be >= 0");
  if (!z || S <= 0 || ldz < N) return fail(-5, "z / S / ldz invalid");
  if (!f || ldf < N) return fail(-8, "f / ldf invalid");
  if (!workspace || workspace_bytes < pgp_prior_sample_workspace_size(N))
    return fail(-10, "workspace too small: %lld < %lld", (long long)workspace_bytes,
                (long long)pgp_prior_sample_workspace_size(N));
  launch_prior_sample(kernel, X, N, jitter, z, S, ldz, f, ldf, (char*)workspace,
                      (hipStream_t)stream);
  return check_launch();
}

int pgp_trsv(int precision, const void* L, int64_t N, int64_t ld, double* dinv, double* b,
             double* x, int trans, void* stream) {
  if (precision < 0 || precision > 2) return fail(-1, "unknown precision %d", precision);
  if (!L || !b || !x || !dinv) return fail(-2, "NULL pointer");
  if (N <= 0 || N % TILE) return fail(-3, "N must be a positive multiple of 64");
  if (ld < N) return fail(-4, "ld < N");
  hipStream_t s = (hipStream_t)stream;
  launch_dinv_raw(precision, L, N, ld, dinv, s);
  launch_trsv_raw_dinv(precision, L, N, ld, dinv, b, x, trans, s);
  return check_launch();
}

int pgp_trsm(int precision, const void* L, int64_t N, int64_t ld, double* dinv, void* S,
             int64_t M, int64_t ld_s, void* stream) {
  if (precision < 0 || precision > 2) return fail(-1, "unknown precision %d", precision);
  if (!L || !S || !dinv) return fail(-2, "NULL pointer");
  if (N <= 0 || N % TILE) return fail(-3, "N must be a positive multiple of 64");
  if (M <= 0 || M % TILE) return fail(-7, "M must be a positive multiple of 64");
  if (ld < N || ld_s < N) return fail(-4, "ld < N");
  hipStream_t s = (hipStream_t)stream;
  launch_dinv_raw(precision, L, N, ld, dinv, s);
  launch_trsm_raw_dinv(precision, L, N, ld, dinv, S, M, ld_s, 0, s);
  return check_launch();
}

int64_t pgp_workspace_size(int64_t N, int32_t D, int32_t precision) {
  (void)D;
  if (N <= 0 || precision < 0 || precision > 2) return -1;
  return ws_layout(N, precision).total;
}

int64_t pgp_predict_workspace_size(int64_t N, int64_t M, int32_t precision) {
  if (N <= 0 || M <= 0 || precision < 0 || precision > 2) return -1;
  return pred_layout(round_up(N, TILE), M, precision).total;
}

int64_t pgp_grad_workspace_size(int64_t N, int32_t D) {
  (void)D;
  if (N <= 0) return -1;
  // sized for the widest storage so any precision fits
  return grad_layout(round_up(N, TILE), PGP_PREC_FP64).total;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// HIP graphs of the single-problem Newton iteration.  An iteration's launches
// depend only on the workspace, the padded size, the factor precision, whether
// B is re-factored and the refinement count (the parameters live in the
// device descriptor), so one captured graph serves every later iteration with
// the same key, in this fit and in later fits on the same workspace
// (optimisation loops, sweeps).  Replays are valid as they are: every
// wavefront launch of an iteration has its own epoch, a member's progress
// word is written by at least two of them per iteration (the two triangular
// sweeps; a factor whose column launches wait on it has two or more of
// those), so no replay finds its own epoch left from the previous one, and
// the ticket counters reset themselves at each launch's end.  Capturing and
// instantiating costs more than launching once, so a key is captured only
// when it recurs in a later fit; a one-off fit runs its launches directly.
// Off with PGP_FIT_NO_GRAPH in the options' flags (or PGP_GRAPH=0), and
// whenever the launch timer is on.
// ---------------------------------------------------------------------------
struct GraphKey {
  int dev;
  const char* ws;
  int Np, prec, refactor, refine, first;
  pgp_tune_entry_t te;  // the schedule the factor's launches follow
  bool operator==(const GraphKey& o) const {
    return dev == o.dev && ws == o.ws && Np == o.Np && prec == o.prec &&
           refactor == o.refactor && refine == o.refine && first == o.first &&
           te.superstep == o.te.superstep && te.column_max == o.te.column_max &&
           te.fuse_max == o.te.fuse_max && te.column_merge == o.te.column_merge;
  }
};
struct GraphEntry {
  GraphKey key;
  int fit_seen;            // the last fit that used the key
  int uses;                // fits that used the key
  int busy;                // iterations holding the graph (handed out, not yet synchronised)
  hipGraphExec_t exec;     // captured once the key recurs in a later fit
};
constexpr size_t GRAPH_CACHE_MAX = 64;
struct GraphCache {
  std::mutex mu;
  std::vector<GraphEntry> e;  // reserved to GRAPH_CACHE_MAX: entries never move
  hipStream_t capture[64] = {};
  int fit_counter = 0;
  GraphCache() { e.reserve(GRAPH_CACHE_MAX); }
};
static GraphCache& graph_cache() {
  static GraphCache c;
  return c;
}
static std::atomic<int64_t> g_graphs_captured{0}, g_graph_launches{0}, g_direct_iters{0};


static bool graphs_enabled(int flags) {
  static const bool env = [] {
    const char* e = getenv("PGP_GRAPH");
    return !(e && e[0] == '0');
  }();
  if (!env || (flags & PGP_FIT_NO_GRAPH)) return false;
  for (int c = 0; c < TC_COUNT; ++c)
    if (timer_on(c)) return false;
  return true;
}

// One iteration's hold on a cache entry: the entry is not evicted while an
// iteration holds it.  release() follows the synchronisation of the stream
// the graph was launched on; a hold dropped without it (an error return)
// synchronises that stream first, so an evicted graph has always finished.
struct GraphHold {
  GraphEntry* ent = nullptr;
  hipStream_t s = nullptr;
  void release() {
    if (!ent) return;
    std::lock_guard<std::mutex> lock(graph_cache().mu);
    --ent->busy;
    ent = nullptr;
  }
  ~GraphHold() {
    if (ent) (void)hipStreamSynchronize(s);
    release();
  }
};

// The graph of key k for fit `fit` (captured now if the key was used by an
// earlier fit), or nullptr: run the launches directly.  A returned graph is
// held by `hold` until the caller has synchronised stream s.  A full cache
// makes room by evicting the entry least recently used by a fit among those
// no iteration holds (keys of workspaces since freed age out this way).
static hipGraphExec_t newton_graph(const GraphKey& k, int fit, const Batch& b, int prec,
                                   int refine, double* gather, hipStream_t s,
                                   GraphHold& hold) {
  GraphCache& gc = graph_cache();
  std::lock_guard<std::mutex> lock(gc.mu);
  GraphEntry* ent = nullptr;
  for (GraphEntry& g : gc.e)
    if (g.key == k) ent = &g;
  if (!ent) {
    if (gc.e.size() < GRAPH_CACHE_MAX) {
      gc.e.push_back(GraphEntry{k, fit, 1, 0, nullptr});
      return nullptr;
    }
    GraphEntry* old = nullptr;
    for (GraphEntry& g : gc.e)
      if (g.busy == 0 && (!old || g.fit_seen < old->fit_seen)) old = &g;
    if (!old) return nullptr;  // every entry held: launch directly
    if (old->exec) (void)hipGraphExecDestroy(old->exec);
    *old = GraphEntry{k, fit, 1, 0, nullptr};
    return nullptr;
  }
  if (ent->fit_seen != fit) {
    ent->fit_seen = fit;
    ++ent->uses;
  }
  if (!ent->exec && ent->uses >= 2 && k.dev >= 0 && k.dev < 64) {
    if (!gc.capture[k.dev] &&
        hipStreamCreateWithFlags(&gc.capture[k.dev], hipStreamNonBlocking) != hipSuccess)
      return nullptr;
    hipStream_t cs = gc.capture[k.dev];
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
    newton_iter(b, k.first ? nullptr : &b, k.refactor ? &b : nullptr, nullptr,
                k.refactor ? nullptr : &b, prec, refine, gather, cs);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(cs, &g);
    hipGraphExec_t exec = nullptr;
    if (e == hipSuccess && g && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) == hipSuccess) {
      ent->exec = exec;
      ++g_graphs_captured;
    }
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();  // a failed capture leaves the launches to run directly
  }
  if (ent->exec) {
    ++ent->busy;
    hold.ent = ent;
    hold.s = s;
  }
  return ent->exec;
}

// Argument checks of the Laplace fit entry points (the workspace aside).
static int check_fit_args(const pgp_problem_t* problem, const pgp_kernel_t* kernel,
                          const pgp_likelihood_t* lik, const pgp_fit_options_t* opts,
                          const pgp_fit_result_t* result) {
  int rc = validate(problem, kernel, lik);
  if (rc) return rc;
  if (!lik) return fail(-3, "likelihood is NULL");
  if (!opts) return fail(-4, "options are NULL");
  if (!result) return fail(-5, "result is NULL");
  if (opts->precision < 0 || opts->precision > 2)
    return fail(-4, "unknown precision %d", opts->precision);
  return 0;
}

// The Laplace fit (Newton loop, final state, resident factor) in workspace
// `ws`, whose size the caller has checked against ws_layout(N, precision).
static int laplace_fit_ws(const pgp_problem_t* problem, const pgp_kernel_t* kernel,
                          const pgp_likelihood_t* lik, const pgp_fit_options_t* opts,
                          pgp_fit_result_t* result, char* ws, hipStream_t s) {
  int rc = 0;
  const int prec = opts->precision;
  const int refine = opts->refine >= 0 ? opts->refine : default_refine(prec);
  Prob p = make_prob(problem, kernel, lik, prec, ws);
  p.state = RS_LAPLACE;
  write_prob(ws, p, 1, s);
  set_self(ws, s);
  const Batch b = single_batch(ws, p.Np);
  double* gather = (double*)(ws + OFF_GATHER);
  launch_gram_self(b, s);
  launch_vec_op(b, OP_ZERO_Z, s);
  rc = check_launch();
  if (rc) return rc;
  memset(result, 0, sizeof *result);
  const bool use_graphs = graphs_enabled(opts->flags);
  int dev = -1;
  if (use_graphs) HIP_OK(hipGetDevice(&dev));
  int fit_id = 0;
  if (use_graphs) {
    std::lock_guard<std::mutex> lock(graph_cache().mu);
    fit_id = ++graph_cache().fit_counter;
  }
  int it = 0;
  double step = INFINITY;
  const bool gauss = lik->kind == PGP_LIK_GAUSSIAN;
  int cur = prec;  // BF16 escalates to FP32 (same storage) if its factor fails
  bool force_factor = false;
  NewtonTrack track;
  while (!newton_done(step, opts->tolerance, track) && it < opts->maxiter) {
    const bool refactor = refactor_now(cur, gauss, it, force_factor, track);
    const bool chord = chord_step(cur, gauss, refactor);
    force_factor = false;
    hipGraphExec_t gexec = nullptr;
    GraphHold hold;
    if (use_graphs)
      gexec = newton_graph(GraphKey{dev, ws, (int)p.Np, cur, refactor ? 1 : 0, refine,
                                    it == 0 ? 1 : 0, tune_get(cur, tune_class_of((int)p.Np))},
                           fit_id, b, cur, refine, gather, s, hold);
    if (gexec) {
      HIP_OK(hipGraphLaunch(gexec, s));
      ++g_graph_launches;
    } else {
      ++g_direct_iters;
      newton_iter(b, it == 0 ? nullptr : &b, refactor ? &b : nullptr, nullptr,
                  refactor ? nullptr : &b, cur, refine, gather, s);
    }
    rc = check_launch();
    if (rc) return rc;
    double h2;
    int32_t inf0;
    HIP_OK(hipMemcpyAsync(&h2, gather, sizeof h2, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(&inf0, ws + OFF_INFO, sizeof inf0, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    hold.release();
    const double st = sqrt(h2);
    if (cur == PGP_PREC_BF16 && ((inf0 & PGP_INFO_NOT_PD) || !(st == st))) {
      // the step was not taken (OP_ZUPD); redo the iteration with FP32 updates
      cur = PGP_PREC_FP32;
      result->status |= PGP_INFO_ESCALATED;
      force_factor = true;
      launch_clear_not_pd(b, s);
      continue;
    }
    if (cur == PGP_PREC_BF16 && stalled(it + 1, st, step)) {
      // the inexact Newton stalls: go on with FP32 updates
      cur = PGP_PREC_FP32;
      result->status |= PGP_INFO_ESCALATED;
      force_factor = true;
    }
    step = st;
    track_step(track, st, chord);
    ++it;
    if (!(step == step) || (inf0 & PGP_INFO_NOT_PD)) break;  // stop, report through status
  }
  result->iterations = it;
  result->step_norm = step;
  if (step > opts->tolerance) result->status |= PGP_INFO_MAXITER;
  if (prec == PGP_PREC_BF16)  // the final factor with FP32 updates (see final_fp32)
    laplace_finalize(b, nullptr, &b, PGP_PREC_FP32, gather, s);
  else
    laplace_finalize(b, gauss ? nullptr : &b, nullptr, cur, gather, s);
  rc = check_launch();
  if (rc) return rc;
  double g4[4];
  int32_t info[4];
  HIP_OK(hipMemcpyAsync(g4, gather, sizeof g4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(info, ws + OFF_INFO, sizeof info, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  fill_result_final(result, g4, info[0]);
  return 0;
}

extern "C" {

int pgp_fit_graph_stats(int64_t* captured, int64_t* graph_iterations, int64_t* direct_iterations) {
  if (!captured || !graph_iterations || !direct_iterations) return fail(-1, "NULL argument");
  *captured = g_graphs_captured.load();
  *graph_iterations = g_graph_launches.load();
  *direct_iterations = g_direct_iters.load();
  return 0;
}

int pgp_laplace_fit(const pgp_problem_t* problem, const pgp_kernel_t* kernel,
                    const pgp_likelihood_t* lik, const pgp_fit_options_t* opts,
                    pgp_fit_result_t* result, void* workspace, int64_t workspace_bytes,
                    void* stream) {
  int rc = check_fit_args(problem, kernel, lik, opts, result);
  if (rc) return rc;
  if (!workspace) return fail(-6, "workspace is NULL");
  const WsLayout L = ws_layout(problem->N, opts->precision);
  if (workspace_bytes < L.total)
    return fail(-7, "workspace too small: %lld < %lld", (long long)workspace_bytes,
                (long long)L.total);
  return laplace_fit_ws(problem, kernel, lik, opts, result, (char*)workspace,
                        (hipStream_t)stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Fit without a caller workspace: the workspace of pgp_laplace_fit_noalloc
// is this file-scope device array, one copy per GPU the library runs on,
// held until the process ends.  It is sized for the largest problem the
// entry point takes (padded N <= PGP_NOALLOC_MAX_NP at any precision).  In
// host code `scratch` names a host-side stand-in, not device memory: its
// device address on the current GPU comes from hipGetSymbolAddress, once per
// device, and only that address is ever handed to kernels or copies.
// ---------------------------------------------------------------------------
static_assert(PGP_NOALLOC_MAX_NP % 64 == 0, "the limit is a whole number of tiles");
constexpr int64_t NOALLOC_BYTES = HEADER + 2 * (int64_t)PGP_NOALLOC_MAX_NP * PGP_NOALLOC_MAX_NP * 8 +
                                  2 * (int64_t)PGP_NOALLOC_MAX_NP * 8 +
                                  (int64_t)PGP_NOALLOC_MAX_NP * TILE * 8 +
                                  (int64_t)NVEC * PGP_NOALLOC_MAX_NP * 8 +
                                  (int64_t)PGP_NOALLOC_MAX_NP * (PGP_NOALLOC_MAX_NP / TILE) * 8 +
                                  csync_words(PGP_NOALLOC_MAX_NP) * 4 + 9 * 256;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES];

// Device address of `scratch` on the current device (looked up on the
// device's first call, with that device current).
static int scratch_address(char** out) {
  int dev = -1;
  HIP_OK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return fail(-100, "device %d beyond the no-allocation table", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(-100, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = (char*)g_scratch_ptr[dev];
  return 0;
}

extern "C" {

int64_t pgp_noalloc_scratch_bytes(void) { return NOALLOC_BYTES; }

int pgp_laplace_fit_noalloc(const pgp_problem_t* problem, const pgp_kernel_t* kernel,
                            const pgp_likelihood_t* lik, const pgp_fit_options_t* opts,
                            pgp_fit_result_t* result, void* stream) {
  int rc = check_fit_args(problem, kernel, lik, opts, result);
  if (rc) return rc;
  const WsLayout L = ws_layout(problem->N, opts->precision);
  if (L.Np > PGP_NOALLOC_MAX_NP || L.total > NOALLOC_BYTES)
    return fail(-1, "N = %lld exceeds the no-allocation limit (N <= %d)",
                (long long)problem->N, PGP_NOALLOC_MAX_NP);
  char* ws = nullptr;
  rc = scratch_address(&ws);
  if (rc) return rc;
  return laplace_fit_ws(problem, kernel, lik, opts, result, ws, (hipStream_t)stream);
}

int pgp_vb_map(const pgp_problem_t* problem, const pgp_kernel_t* kernel,
               const pgp_likelihood_t* lik, int32_t precision, const double* weight,
               void* workspace, int64_t workspace_bytes, void* stream) {
  int rc = validate(problem, kernel, lik);
  if (rc) return rc;
  if (!lik) return fail(-3, "likelihood is NULL");
  if (precision < 0 || precision > 2) return fail(-4, "unknown precision %d", precision);
  if (!weight) return fail(-5, "weight is NULL");
  if (!problem->weight) return fail(-1, "problem->weight (the output) is NULL");
  const WsLayout L = ws_layout(problem->N, precision);
  if (!workspace || workspace_bytes < L.total) return fail(-6, "workspace too small");
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  Prob p = make_prob(problem, kernel, lik, precision, ws);
  p.out_prec = nullptr;
  p.out_latent = nullptr;
  p.state = RS_VBMAP;
  write_prob(ws, p, 1, s);
  set_self(ws, s);
  const Batch b = single_batch(ws, p.Np);
  launch_gram_self(b, s);
  launch_load_state(b, weight, nullptr, s);
  launch_gemv_sym(b, V_Z, V_F, 0, s);  // f = K w
  launch_lik_final(b, 0, s);           // d1 at f
  launch_vec_op(b, OP_VB_RHS, s);      // X1 = f + sigma d1
  launch_copy_slot(b, V_X1, V_T, s);
  factor(b, precision, 1, s);          // sigma^2 I + K
  solve_B(b, precision, s);            // U = (sigma^2 I + K)^-1 X1
  launch_copy_slot(b, V_U, V_Z, s);
  launch_vec_op(b, OP_OUT, s);         // problem->weight <- U
  return check_launch();
}

int pgp_laplace_set_state(const pgp_problem_t* problem, const pgp_kernel_t* kernel,
                          const pgp_likelihood_t* lik, int32_t precision, const double* weight,
                          const double* W, pgp_fit_result_t* result, void* workspace,
                          int64_t workspace_bytes, void* stream) {
  int rc = validate(problem, kernel, lik);
  if (rc) return rc;
  if (!lik) return fail(-3, "likelihood is NULL");
  if (precision < 0 || precision > 2) return fail(-4, "unknown precision %d", precision);
  if (!weight) return fail(-5, "weight is NULL");
  const WsLayout L = ws_layout(problem->N, precision);
  if (!workspace || workspace_bytes < L.total) return fail(-9, "workspace too small");
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  Prob p = make_prob(problem, kernel, lik, precision, ws);
  p.state = RS_LAPLACE;
  write_prob(ws, p, 1, s);
  set_self(ws, s);
  const Batch b = single_batch(ws, p.Np);
  launch_gram_self(b, s);
  // weight -> Z, f = K a, log p at f, then W and S = sqrt(max(W, 0)) from the caller
  launch_load_state(b, weight, nullptr, s);
  launch_gemv_sym(b, V_Z, V_F, 0, s);
  launch_lik_final(b, 0, s);  // W = -d2 log p(K a), S = sqrt(max(W, 0))
  if (W) launch_load_state(b, nullptr, W, s);  // caller-given precision instead
  factor(b, precision, 0, s);
  double* gather = (double*)(ws + OFF_GATHER);
  launch_reduce(b, RED_FINAL, gather, s);
  rc = check_launch();
  if (rc) return rc;
  double g4[4];
  int32_t info[4];
  HIP_OK(hipMemcpyAsync(g4, gather, sizeof g4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(info, ws + OFF_INFO, sizeof info, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  if (result) {
    memset(result, 0, sizeof *result);
    fill_result_final(result, g4, info[0]);
  }
  launch_vec_op(b, OP_OUT, s);
  return check_launch();
}

int pgp_predict(const pgp_kernel_t* kernel, const double* Xtest, int64_t M, double* mean,
                double* var, void* workspace, int64_t workspace_bytes, void* predict_workspace,
                int64_t predict_workspace_bytes, void* stream) {
  if (!kernel) return fail(-1, "kernel is NULL");
  if (!Xtest || M <= 0) return fail(-2, "Xtest / M invalid");
  if (!mean || !var) return fail(-4, "mean / var are NULL");
  if (!workspace) return fail(-6, "workspace is NULL");
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  Prob p;
  int rc = read_prob(ws, p, s);
  if (rc) return rc;
  rc = check_resident(p, RS_PREDICT, "pgp_predict", kernel, -1, nullptr, 0);
  if (rc) return rc;
  const WsLayout L = ws_layout(p.N, p.prec);
  if (workspace_bytes < L.total) return fail(-7, "workspace too small");
  const PredLayout PL = pred_layout(p.Np, M, p.prec);
  if (!predict_workspace || predict_workspace_bytes < PL.total)
    return fail(-9, "predict workspace too small: %lld < %lld",
                (long long)predict_workspace_bytes, (long long)PL.total);
  char* pw = (char*)predict_workspace;
  p.Xt = Xtest;
  p.M = (int32_t)M;
  p.Mp = (int32_t)PL.Mp;
  p.S = pw + PL.S;
  p.mupart = (double*)(pw + PL.mupart);
  p.qpart = predict_refine(p.prec) ? (double*)(pw + PL.qpart) : nullptr;
  p.pmean = mean;
  p.pvar = var;
  write_prob(ws, p, 0, s);
  const Batch b = single_batch(ws, p.Np);
  predict_launches(b, p.prec, (int)M, s);
  rc = check_launch();
  if (rc) return rc;
  int bits = 0;
  rc = read_info_bits(ws, s, &bits);
  return rc ? rc : bits;
}

int pgp_predict_covariance(const pgp_kernel_t* kernel, const double* Xtest, int64_t M,
                           double* cov, int64_t ld, void* workspace, int64_t workspace_bytes,
                           void* predict_workspace, int64_t predict_workspace_bytes,
                           void* stream) {
  if (!kernel) return fail(-1, "kernel is NULL");
  if (!Xtest || M <= 0) return fail(-2, "Xtest / M invalid");
  if (!cov) return fail(-4, "cov is NULL");
  if (ld < M) return fail(-5, "ld < M");
  if (!workspace) return fail(-6, "workspace is NULL");
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  Prob p;
  int rc = read_prob(ws, p, s);
  if (rc) return rc;
  rc = check_resident(p, RS_PREDICT, "pgp_predict_covariance", kernel, -1, nullptr, 0);
  if (rc) return rc;
  const WsLayout L = ws_layout(p.N, p.prec);
  if (workspace_bytes < L.total) return fail(-7, "workspace too small");
  const PredLayout PL = pred_layout(p.Np, M, p.prec);
  if (!predict_workspace || predict_workspace_bytes < PL.total)
    return fail(-9, "predict workspace too small: %lld < %lld",
                (long long)predict_workspace_bytes, (long long)PL.total);
  char* pw = (char*)predict_workspace;
  p.Xt = Xtest;
  p.M = (int32_t)M;
  p.Mp = (int32_t)PL.Mp;
  p.S = pw + PL.S;
  p.mupart = (double*)(pw + PL.mupart);
  p.qpart = nullptr;
  p.pmean = nullptr;
  p.pvar = nullptr;
  write_prob(ws, p, 0, s);
  const Batch b = single_batch(ws, p.Np);
  launch_cross_prep(b, p.prec, (int)M, s);
  launch_trsm(b, p.prec, (int)PL.Mp, 0, s);
  launch_pred_cov(b, p.prec, (int)PL.Mp, cov, ld, s);
  rc = check_launch();
  if (rc) return rc;
  int bits = 0;
  rc = read_info_bits(ws, s, &bits);
  return rc ? rc : bits;
}

int pgp_laplace_grad(const pgp_kernel_t* kernel, const pgp_likelihood_t* lik, double* grad_out,
                     void* workspace, int64_t workspace_bytes, void* grad_workspace,
                     int64_t grad_workspace_bytes, void* stream) {
  if (!kernel) return fail(-1, "kernel is NULL");
  if (!lik) return fail(-2, "likelihood is NULL");
  if (!grad_out) return fail(-3, "grad_out is NULL");
  if (!workspace) return fail(-4, "workspace is NULL");
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  Prob p;
  int rc = read_prob(ws, p, s);
  if (rc) return rc;
  rc = check_resident(p, 1u << RS_LAPLACE, "pgp_laplace_grad", kernel, -1, lik, -2);
  if (rc) return rc;
  if (p.D > grad_maxd()) return fail(-1, "gradient supports D <= %d (got %d)", grad_maxd(), p.D);
  if (p.lik_kind == PGP_LIK_PROBIT_CLOSED)
    return fail(-2, "gradient of the closed-form probit variant is not provided");
  const WsLayout L = ws_layout(p.N, p.prec);
  if (workspace_bytes < L.total) return fail(-5, "workspace too small");
  const GradLayout G = grad_layout(p.Np, PGP_PREC_FP64);
  if (!grad_workspace || grad_workspace_bytes < G.total)
    return fail(-7, "gradient workspace too small: %lld < %lld", (long long)grad_workspace_bytes,
                (long long)G.total);
  char* gw = (char*)grad_workspace;
  p.C = (double*)(gw + G.C);
  p.S = gw + G.G2;
  p.Mp = p.Np;
  write_prob(ws, p, 0, s);
  const Batch b = single_batch(ws, p.Np);
  const int prec = p.prec;
  launch_lik_final(b, 1, s);  // G, W, S, LL, D3 at the resident f
  // C = S Linv^T Linv S
  launch_fill_rhs(b, prec, 1, s);
  launch_trsm(b, prec, p.Np, 1, s);
  launch_sym_from_inv(b, prec, 1, s);
  // Sigma_ii from V = L^-1 (S K)
  launch_fill_rhs(b, prec, 0, s);
  launch_trsm(b, prec, p.Np, 0, s);
  launch_sigma_diag(b, prec, s);
  // shat = s2 - C K s2, s2 = -1/2 Sigma d3;  X2 <- K shat
  launch_vec_op(b, OP_GRAD_S2, s);       // X2 = -0.5 X1 .* D3
  launch_gemv_sym(b, V_X2, V_X4, 0, s);  // X4 = K s2
  launch_gemv_sym(b, V_X4, V_T, 10, s);  // T = C X4   (mode 10: matrix C)
  launch_vec_op(b, OP_GRAD_SHAT, s);     // X3 = X2 - T
  launch_gemv_sym(b, V_X3, V_X2, 0, s);  // X2 = K shat
  double* lik_out = (double*)(gw + G.out);
  double* kern_out = lik_out + 64;
  launch_lik_grad_terms(b, lik_out, s);
  launch_grad_trace(b, (double*)(gw + G.partials), kern_out, s);
  rc = check_launch();
  if (rc) return rc;
  std::vector<double> hl(1 + PGP_MAX_CUT), hk(grad_npar());
  HIP_OK(hipMemcpyAsync(hl.data(), lik_out, hl.size() * 8, hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(hk.data(), kern_out, hk.size() * 8, hipMemcpyDeviceToHost, s));
  int bits = 0;
  rc = read_info_bits(ws, s, &bits);  // also waits for the copies
  if (rc) return rc;
  const int D = p.D, gm = grad_maxd();
  // an invalid factor or an expired wait in the adjoint's solves: NaN
  const double bad = bits ? NAN : 0.0;
  int o = 0;
  grad_out[o++] = hk[0] + bad;
  for (int d = 0; d < D; ++d) grad_out[o++] = hk[1 + d] + bad;
  for (int d = 0; d < D; ++d) grad_out[o++] = (p.periodic ? hk[1 + gm + d] : 0.0) + bad;
  grad_out[o++] = hl[0] + bad;
  for (int c = 0; c < PGP_MAX_CUT; ++c) grad_out[o++] = hl[1 + c] + bad;
  return bits;
}

// Gradient of objective_VB at the resident VB fixed point (after pgp_vb_fit).
// With A = sigma^2 I + K, C = A^-1, g(a) = C (K a + sigma d1(K a)) the
// fixed-point map and d2 = -W its curvature, the implicit-function adjoint
// (I - dg/da)^T u = K (a - d1) reduces to u = A v with
//   sigma (sigma I + K W) v = K r,  r = a - d1,
// solved by Woodbury with B_s = sigma I + S K S (S = sqrt W):
//   v = (K r - K S B_s^-1 S K r) / sigma^2.
// Then  d psi / d theta_K = sum_ij dK_ij G_ij,
//   G = C/2 + a a^T/2 + sym((sigma d2 .* v - d1) a^T)
//     = C/2 - a a^T/2 + sym(shat a^T),  shat = a - d1 - sigma W .* v
// (the trace kernel's form; the reference's two trace terms add up to N/2
// and carry no gradient), and for the likelihood parameters
//   d/d theta_l = sum_i [-d_theta log p + sigma v_i d_theta d1_i]
//   (+ for sigma: -N/sigma + sigma tr C + v . (d1 - 2 sigma a)).
// The factor predict uses (B with W = 1/sigma^2) is re-formed afterwards.
int pgp_vb_grad(const pgp_kernel_t* kernel, const pgp_likelihood_t* lik, double* grad_out,
                void* workspace, int64_t workspace_bytes, void* grad_workspace,
                int64_t grad_workspace_bytes, void* stream) {
  if (!kernel) return fail(-1, "kernel is NULL");
  if (!lik) return fail(-2, "likelihood is NULL");
  if (!grad_out) return fail(-3, "grad_out is NULL");
  if (!workspace) return fail(-4, "workspace is NULL");
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  Prob p;
  int rc = read_prob(ws, p, s);
  if (rc) return rc;
  rc = check_resident(p, 1u << RS_VB, "pgp_vb_grad", kernel, -1, lik, -2);
  if (rc) return rc;
  if (p.D > grad_maxd()) return fail(-1, "gradient supports D <= %d (got %d)", grad_maxd(), p.D);
  if (p.lik_kind == PGP_LIK_PROBIT_CLOSED)
    return fail(-2, "gradient of the closed-form probit variant is not provided");
  const WsLayout L = ws_layout(p.N, p.prec);
  if (workspace_bytes < L.total) return fail(-5, "workspace too small");
  const GradLayout G = grad_layout(p.Np, PGP_PREC_FP64);
  if (!grad_workspace || grad_workspace_bytes < G.total)
    return fail(-7, "gradient workspace too small: %lld < %lld", (long long)grad_workspace_bytes,
                (long long)G.total);
  char* gw = (char*)grad_workspace;
  p.C = (double*)(gw + G.C);
  p.S = gw + G.G2;
  p.Mp = p.Np;
  write_prob(ws, p, 0, s);
  const Batch b = single_batch(ws, p.Np);
  const int prec = p.prec;
  double* gather = (double*)(ws + OFF_GATHER);
  // C = A^-1 = S B^-1 S from the resident factor (W = 1/sigma^2, S = 1/sigma)
  launch_vec_op(b, OP_VB_PREC, s);
  launch_fill_rhs(b, prec, 1, s);
  launch_trsm(b, prec, p.Np, 1, s);
  launch_sym_from_inv(b, prec, 1, s);
  // f = K a; d1, W = -d2, S = sqrt(max(W, 0)) of the likelihood
  launch_gemv_sym(b, V_Z, V_F, 0, s);
  launch_lik_final(b, 0, s);
  launch_reduce(b, RED_FINAL, gather, s);  // [3]: count of W < 0
  // v = (K r - K S B_s^-1 S K r) / sigma^2
  launch_vec_op(b, OP_VBG_R, s);         // X1 = r
  launch_gemv_sym(b, V_X1, V_X2, 0, s);  // X2 = K r
  launch_vec_op(b, OP_VBG_T, s);         // T = S K r
  factor(b, prec, 2, s);                 // B_s
  solve_B(b, prec, s);                   // U = B_s^-1 T
  launch_vec_op(b, OP_VBG_Z, s);         // X4 = S U
  launch_gemv_sym(b, V_X4, V_X3, 0, s);  // X3 = K S U
  launch_vec_op(b, OP_VBG_V, s);         // X4 = v
  launch_vec_op(b, OP_VBG_SHAT, s);      // X3 = shat, X1 = 0, X2 = sigma v
  double* lik_out = (double*)(gw + G.out);
  double* kern_out = lik_out + 64;
  double* sig_out = kern_out + 64;
  launch_lik_grad_terms(b, lik_out, s);
  launch_vb_sigma_terms(b, sig_out, s);
  launch_grad_trace(b, (double*)(gw + G.partials), kern_out, s);
  // the resident state of the fit again: W = 1/sigma^2 and its B factor
  launch_vec_op(b, OP_VB_PREC, s);
  factor(b, prec, 0, s);
  rc = check_launch();
  if (rc) return rc;
  std::vector<double> hl(1 + PGP_MAX_CUT), hk(grad_npar());
  double hs[2], g4[4];
  HIP_OK(hipMemcpyAsync(hl.data(), lik_out, hl.size() * 8, hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(hk.data(), kern_out, hk.size() * 8, hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(hs, sig_out, sizeof hs, hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(g4, gather, sizeof g4, hipMemcpyDeviceToHost, s));
  int bits = 0;
  rc = read_info_bits(ws, s, &bits);  // also waits for the copies
  if (rc) return rc;
  const int D = p.D, gm = grad_maxd();
  const double sigma = p.noise_std;
  // W < 0 somewhere (the probit floor): the Woodbury form does not hold;
  // an invalid factor or an expired wait in the adjoint's solves: NaN too
  const double bad = (g4[3] > 0 || bits) ? NAN : 0.0;
  int o = 0;
  grad_out[o++] = hk[0] + bad;
  for (int d = 0; d < D; ++d) grad_out[o++] = hk[1 + d] + bad;
  for (int d = 0; d < D; ++d) grad_out[o++] = (p.periodic ? hk[1 + gm + d] : 0.0) + bad;
  grad_out[o++] = hl[0] - p.N / sigma + sigma * hs[0] + hs[1] + bad;
  for (int c = 0; c < PGP_MAX_CUT; ++c) grad_out[o++] = hl[1 + c] + bad;
  return bits;
}

int pgp_vb_fit(const pgp_problem_t* problem, const pgp_kernel_t* kernel,
               const pgp_likelihood_t* lik, const pgp_fit_options_t* opts,
               pgp_fit_result_t* result, void* workspace, int64_t workspace_bytes,
               void* grad_workspace, int64_t grad_workspace_bytes, void* stream) {
  int rc = validate(problem, kernel, lik);
  if (rc) return rc;
  if (!lik) return fail(-3, "likelihood is NULL");
  if (!opts || !result) return fail(-4, "options / result are NULL");
  const int prec = opts->precision;
  if (prec < 0 || prec > 2) return fail(-4, "unknown precision %d", prec);
  const WsLayout L = ws_layout(problem->N, prec);
  if (!workspace || workspace_bytes < L.total) return fail(-7, "workspace too small");
  const GradLayout G = grad_layout(L.Np, PGP_PREC_FP64);
  if (!grad_workspace || grad_workspace_bytes < G.total)
    return fail(-9, "gradient workspace too small (needed for objective_VB)");
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  char* gw = (char*)grad_workspace;
  const int refine = opts->refine >= 0 ? opts->refine : default_refine(prec);
  Prob p = make_prob(problem, kernel, lik, prec, ws);
  p.C = (double*)(gw + G.C);
  p.S = gw + G.G2;
  p.Mp = p.Np;
  p.state = RS_VB;
  write_prob(ws, p, 1, s);
  set_self(ws, s);
  const Batch b = single_batch(ws, p.Np);
  double* gather = (double*)(ws + OFF_GATHER);
  launch_gram_self(b, s);
  factor(b, prec, 1, s);  // A = sigma^2 I + K
  launch_vec_op(b, OP_ZERO_Z, s);
  rc = check_launch();
  if (rc) return rc;
  memset(result, 0, sizeof *result);
  int it = 0;
  double step = INFINITY;
  while (step > opts->tolerance && it < opts->maxiter) {
    launch_gemv_sym(b, V_Z, V_F, 0, s);
    launch_lik_final(b, 0, s);
    launch_vec_op(b, OP_VB_RHS, s);
    launch_copy_slot(b, V_X1, V_T, s);
    solve_B(b, prec, s);
    for (int r = 0; r < refine; ++r) {
      launch_gemv_sym(b, V_U, V_Q, 3, s);
      launch_copy_slot(b, V_Q, V_T, s);
      launch_trsv(b, prec, 0, V_T, V_X3, s);
      launch_trsv(b, prec, 1, V_X3, V_X4, s);
      launch_vec_op(b, OP_VB_IRCORR, s);
    }
    launch_vec_op(b, OP_VB_STEP, s);
    launch_vec_op(b, OP_ZUPD, s);
    launch_reduce(b, RED_STEP, gather, s);
    rc = check_launch();
    if (rc) return rc;
    double h2;
    HIP_OK(hipMemcpyAsync(&h2, gather, sizeof h2, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    step = sqrt(h2);
    ++it;
    if (!(step == step)) break;
  }
  result->iterations = it;
  result->step_norm = step;
  if (step > opts->tolerance) result->status |= PGP_INFO_MAXITER;
  // objective_VB pieces with the factor of A still resident
  launch_gemv_sym(b, V_Z, V_F, 0, s);
  launch_lik_final(b, 0, s);
  launch_reduce(b, RED_FINAL, gather, s);
  launch_fill_rhs(b, prec, 1, s);
  launch_trsm(b, prec, p.Np, 1, s);
  launch_sym_from_inv(b, prec, 0, s);
  double* tr_out = (double*)(gw + G.out);
  launch_vb_trace(b, tr_out, s);
  // precision = 1/sigma^2 and the B-form factor predict uses
  launch_vec_op(b, OP_VB_PREC, s);
  factor(b, prec, 0, s);
  launch_vec_op(b, OP_OUT, s);
  rc = check_launch();
  if (rc) return rc;
  double g4[4], tr[2];
  int32_t info[4];
  HIP_OK(hipMemcpyAsync(g4, gather, sizeof g4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(tr, tr_out, sizeof tr, hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(info, ws + OFF_INFO, sizeof info, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  const double sigma = lik->noise_std;
  const double N = (double)problem->N;
  result->sum_log_lik = g4[0];
  result->half_f_dot_a = 0.5 * g4[1];
  result->log_det = 2.0 * g4[2];  // log |sigma^2 I + K|
  // objective_VB (probit/implicit/VB.py:33-39)
  result->objective = 0.5 * tr[1] + 0.5 * sigma * sigma * tr[0] + 0.5 * g4[1] -
                      N * log(sigma) + g4[2] - 0.5 * N - g4[0];
  apply_info(result, info[0]);  // NaN objective on a non-positive pivot or an expired wait
  return 0;
}

}  // extern "C"

// ===========================================================================
// Continuous batching: many fits share every launch of a Newton iteration.
//
// A request joins at the next step after submit(); each step runs one
// Newton iteration for every request in flight (one launch per blocked
// Cholesky / solve step for all of them together), then finalises and
// predicts the requests that converged, which leave the batch.  Every
// member's arithmetic is independent of the other members (per-member
// tiles, fixed-order per-member reductions), so a request's result is the
// same alone and in any mix.
//
// Asynchronous form: the members are split by size into two lanes, each with
// its own streams and its own step cadence.  A step's duration is set by
// its largest members' factorisations, so the lane of small members runs
// several steps while the lane of large members runs one, and fills the GPU
// during the large members' latency-bound column steps.
// ===========================================================================
namespace pgp {

// Device pointer arrays of a lane: 0 admitted / flag clearing, 1 Newton
// members, 2 their factored ones, 3 finalised members, 4 their factored
// ones, 5 predicted ones, 6 and 7 the FP32-escalated factored members of 2
// and 4, 8 the Newton members not re-factored.  Arrays 3, 4, 5 and 7 belong
// to the lane's finalisation stream.
constexpr int BATCH_ARRAYS = 10;
constexpr int NLANES = 4;

struct BItem {
  pgp_batch_item_t* item;
  int64_t tag;
  char* ws;
  Prob p;
  int phase;  // 0 = new, 1 = Newton, 2 = converged (finalise next), 3 = finalising
  int iter;
  int lane;   // asynchronous form: size class (0 = smallest)
  double step;
  bool gauss;
  bool done;
  bool esc = false;    // BF16 batcher: factored with FP32 updates from now on
  bool clear = false;  // its NOT_PD flag is to be cleared before the next step
  bool refac = false;  // re-factor at the next step (escalated Gaussian member)
  double prev = INFINITY;  // step norm of the previous iteration
  NewtonTrack track;       // re-factor rule (refactor_now)
  bool chord = false;      // this member's current iteration is a chord step
};

struct Lane {
  hipStream_t s = nullptr;   // Newton steps (the lane's own stream)
  hipStream_t s2 = nullptr;  // finalisation and predict, beside the Newton step
  bool own_s = false;
  hipEvent_t ev = nullptr, ev_fork = nullptr, ev_join = nullptr;
  Prob** arr[BATCH_ARRAYS];
  std::vector<int> hnp[BATCH_ARRAYS], hm[BATCH_ARRAYS];
  double* gather;      // per-member step norms (device)
  double* gather_fin;  // 8 doubles per finalised member (device)
  uint32_t* ticket;    // start-order tickets of launches on s and on s2
  uint32_t* ticket2;
  Prob* ddesc;             // device staging of descriptors (2 * cap)
  Prob* hdesc = nullptr;   // pinned host staging of descriptors (2 * cap)
  double* hstep = nullptr;  // pinned copies of gather / gather_fin
  double* hfin = nullptr;
  bool launched = false;
  std::vector<BItem*> act_list, fin_list;
};

}  // namespace pgp

struct pgp_batcher {
  pgp_fit_options_t opts;
  int refine;
  int cap;
  hipStream_t caller = nullptr;     // the stream given at creation
  hipEvent_t ev_caller = nullptr;   // its state at each launch / step
  // host-side counters (pgp_batcher_stats): seconds spent enqueueing lane
  // steps and reporting them, seconds finish() waited for the GPU, steps
  double host_enqueue_s = 0.0, host_wait_s = 0.0, host_report_s = 0.0;
  int64_t lane_steps = 0;

  int lane_np[pgp::NLANES - 1];  // size class bounds on Np (0: unused)
  int split_min = 0;  // members in flight from which lanes past 1 take members
  pgp::Lane lane[pgp::NLANES];
  std::list<pgp::BItem> items;  // stable addresses: lanes hold pointers
  std::vector<pgp::BItem> pending;
};

namespace pgp {

static int64_t lane_table_bytes(int cap) {
  return align256((int64_t)BATCH_ARRAYS * cap * 8) + 2 * align256((int64_t)cap * 8 * 8) + 256 +
         align256((int64_t)2 * cap * (int64_t)sizeof(Prob));
}
static int64_t batcher_table_bytes(int cap) { return NLANES * lane_table_bytes(cap); }

static bool on_s2(int which) { return which == 3 || which == 4 || which == 5 || which == 7; }

// Write the descriptors of a batch's members (held in v[i]->p) to their
// workspaces with one staged copy and one kernel.  `region` (0 or 1) picks
// the half of the lane's staging buffers, so that two writes of one step do
// not overlap; `arr` is the batch's device pointer array (already written).
static int write_probs(pgp_batcher* bt, Lane& L, const std::vector<BItem*>& v,
                       Prob* const* arr, int region, int reset, hipStream_t s) {
  if (v.empty()) return 0;
  Prob* h = L.hdesc + (size_t)region * bt->cap;
  Prob* d = L.ddesc + (size_t)region * bt->cap;
  for (size_t i = 0; i < v.size(); ++i) h[i] = v[i]->p;
  HIP_OK(hipMemcpyAsync(d, h, v.size() * sizeof(Prob), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_scatter_probs, dim3((unsigned)v.size()), dim3(256), 0, s, arr, d,
                     (int)v.size(), reset);
  return check_launch();
}

// Members are ordered by padded size, largest first (stable), so that the
// members still active at any block column form a prefix of the batch.
static Batch make_batch(Lane& L, int which, std::vector<BItem*>& v) {
  std::stable_sort(v.begin(), v.end(),
                   [](const BItem* a, const BItem* b) { return a->p.Np > b->p.Np; });
  std::vector<Prob*> ptrs;
  ptrs.reserve(v.size());
  std::vector<int>& hnp = L.hnp[which];
  std::vector<int>& hm = L.hm[which];
  hnp.clear();
  hm.clear();
  int maxNp = 0;
  for (BItem* b : v) {
    ptrs.push_back((Prob*)(b->ws + OFF_PROB));
    hnp.push_back(b->p.Np);
    hm.push_back(b->p.M);
    maxNp = std::max(maxNp, (int)b->p.Np);
  }
  write_ptrs(L.arr[which], ptrs, on_s2(which) ? L.s2 : L.s);
  Batch b;
  b.dprobs = L.arr[which];
  b.count = (int)v.size();
  b.maxNp = maxNp;
  b.hnp = hnp.data();
  b.hm = hm.data();
  b.ticket = on_s2(which) ? L.ticket2 : L.ticket;
  return b;
}

// Move pending requests into the batch up to capacity and give each a lane.
static void admit(pgp_batcher* bt) {
  size_t take = 0;
  while (take < bt->pending.size() && (int)bt->items.size() < bt->cap) {
    BItem b = bt->pending[take++];
    b.lane = 0;
    for (int k = 0; k < NLANES - 1; ++k)
      if (bt->lane_np[k] > 0 && b.p.Np > bt->lane_np[k]) b.lane = k + 1;
    // a lightly loaded engine keeps every large member in lane 1
    if (b.lane > 1 && (int)bt->items.size() < bt->split_min) b.lane = 1;
    bt->items.push_back(b);
  }
  bt->pending.erase(bt->pending.begin(), bt->pending.begin() + take);
}

// New members of `fresh` get their descriptor, Gram matrix and a zero
// weight (on the lane's Newton stream).
static int start_members(pgp_batcher* bt, Lane& L, std::vector<BItem*>& fresh) {
  if (fresh.empty()) return 0;
  const Batch bn = make_batch(L, 0, fresh);
  const int rc = write_probs(bt, L, fresh, bn.dprobs, 0, 1, L.s);
  if (rc) return rc;
  launch_gram_self(bn, L.s);
  launch_vec_op(bn, OP_ZERO_Z, L.s);
  for (BItem* b : fresh) b->phase = 1;
  return check_launch();
}

// Status word of each member into slot 4 of its 8 gathered doubles.
__global__ void k_gather_info(Prob* const* arr, int n, double* gather) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) gather[8 * i + 4] = (double)arr[i]->info[0];
}
static void launch_gather_info(const Batch& b, double* gather, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_info, dim3((b.count + 255) / 256), dim3(256), 0, s, b.dprobs,
                     b.count, gather);
}

// Enqueue on the lane's finalisation stream the final state (f, W, log p,
// B factored at W) and the predict of the members in `conv`; 8 result
// doubles per member land in the lane's gather_fin, copied to its hfin.
static int finalize_and_predict(pgp_batcher* bt, Lane& L, std::vector<BItem*>& conv) {
  hipStream_t s = L.s2;
  const int prec = bt->opts.precision;
  std::vector<BItem*> conv_fac, conv_fac32, pred;
  for (BItem* b : conv)
    if (prec == PGP_PREC_BF16)
      conv_fac32.push_back(b);  // the final factor with FP32 updates (see final_fp32)
    else if (!b->gauss)
      (b->esc ? conv_fac32 : conv_fac).push_back(b);
  const Batch bc = make_batch(L, 3, conv);
  const Batch bcf = make_batch(L, 4, conv_fac);
  const Batch bcf32 = make_batch(L, 7, conv_fac32);
  laplace_finalize(bc, &bcf, &bcf32, prec, L.gather_fin, s);
  int maxM = 0;
  for (BItem* b : conv) {
    pgp_batch_item_t* it = b->item;
    if (!it->Xtest) continue;
    const PredLayout PL = pred_layout(b->p.Np, it->M, prec);
    char* pw = (char*)it->predict_workspace;
    b->p.Xt = it->Xtest;
    b->p.M = (int32_t)it->M;
    b->p.Mp = (int32_t)PL.Mp;
    b->p.S = pw + PL.S;
    b->p.mupart = (double*)(pw + PL.mupart);
    b->p.qpart = predict_refine(prec) ? (double*)(pw + PL.qpart) : nullptr;
    b->p.pmean = it->mean;
    b->p.pvar = it->var;
    pred.push_back(b);
    maxM = std::max(maxM, (int)it->M);
  }
  if (!pred.empty()) {
    const Batch bp = make_batch(L, 5, pred);
    const int rc = write_probs(bt, L, pred, bp.dprobs, 1, 0, s);
    if (rc) return rc;
    predict_launches(bp, prec, maxM, s);
    // the status words again: the predict solve's bounded waits may have run out
    launch_gather_info(bc, L.gather_fin, s);
  }
  HIP_OK(hipMemcpyAsync(L.hfin, L.gather_fin, conv.size() * 64, hipMemcpyDeviceToHost, s));
  return check_launch();
}

// Results of finalised members from their 8 gathered doubles each.
static void report(pgp_batcher* bt, const std::vector<BItem*>& conv, const double* h,
                   int64_t* done_tags, int32_t max_done, int32_t* n_done) {
  for (size_t i = 0; i < conv.size(); ++i) {
    BItem* b = conv[i];
    pgp_fit_result_t* r = &b->item->result;
    const int32_t esc = r->status & PGP_INFO_ESCALATED;
    memset(r, 0, sizeof *r);
    r->status = esc;
    r->iterations = b->iter;
    r->step_norm = b->step;
    if (b->step > bt->opts.tolerance) r->status |= PGP_INFO_MAXITER;
    fill_result_final(r, &h[8 * i], (int)h[8 * i + 4]);
    if (done_tags && *n_done < max_done) done_tags[*n_done] = b->tag;
    ++*n_done;
    b->done = true;
  }
}

// Drop finished members (the others keep their addresses).
static void drop_done(pgp_batcher* bt) {
  bt->items.remove_if([](const BItem& b) { return b.done; });
}

// BF16 batcher, as the single-problem fit: a member whose factor lost
// definiteness (NaN step; its weight was left unchanged) redoes the
// iteration with FP32 updates; one whose inexact Newton stalls (from the
// third iteration, a step norm above half the previous one) continues with
// FP32 updates from its current weight.  Returns true if the iteration is
// to be redone (not counted).
static bool escalate(const pgp_batcher* bt, BItem* b) {
  if (bt->opts.precision != PGP_PREC_BF16 || b->esc) return false;
  const bool failed = !(b->step == b->step);
  if (!failed && !stalled(b->iter + 1, b->step, b->prev)) return false;
  b->esc = true;
  b->refac = true;
  b->clear = failed;
  b->item->result.status |= PGP_INFO_ESCALATED;
  if (failed) b->step = INFINITY;
  return failed;
}

// Clear the non-positive-pivot bit of the members of `v` escalated in their
// previous step (array 0, after the admission's launches on the same stream).
static void clear_escalated(Lane& L, const std::vector<BItem*>& v) {
  std::vector<BItem*> clr;
  for (BItem* b : v)
    if (b->clear) {
      clr.push_back(b);
      b->clear = false;
    }
  if (clr.empty()) return;
  launch_clear_not_pd(make_batch(L, 0, clr), L.s);
}

static bool converged(const pgp_batcher* bt, const BItem* b) {
  return newton_done(b->step, bt->opts.tolerance, b->track) || b->iter >= bt->opts.maxiter ||
         !(b->step == b->step);
}

// Step norm of a member after its Newton iteration; true if it converged.
static bool advance(pgp_batcher* bt, BItem* b, double h2) {
  b->step = sqrt(h2);
  if (escalate(bt, b)) return false;
  b->iter++;
  b->prev = b->step;
  track_step(b->track, b->step, b->chord);
  return converged(bt, b);
}

// Factor lists of one Newton step (refactor_now per member; escalated BF16
// members factor with FP32 updates).
static void split_factor(const pgp_batcher* bt, const std::vector<BItem*>& act,
                         std::vector<BItem*>& fac, std::vector<BItem*>& fac32,
                         std::vector<BItem*>& nofac) {
  const int prec = bt->opts.precision;
  for (BItem* b : act) {
    const int cur = b->esc ? PGP_PREC_FP32 : prec;
    const bool f = refactor_now(cur, b->gauss, b->iter, b->refac, b->track);
    b->chord = chord_step(cur, b->gauss, f);
    if (f)
      (b->esc ? fac32 : fac).push_back(b);
    else
      nofac.push_back(b);
    b->refac = false;
  }
}

// Enqueue one step of lane `li` (finalisation of the members that converged
// in its previous step beside one Newton iteration of the others) and record
// its completion event.  Nothing is enqueued if the lane has no members.
static int lane_launch(pgp_batcher* bt, int li) {
  Lane& L = bt->lane[li];
  hipStream_t s = L.s;
  const int prec = bt->opts.precision;
  L.fin_list.clear();
  L.act_list.clear();
  std::vector<BItem*> fresh;
  for (BItem& b : bt->items) {
    if (b.lane != li) continue;
    if (b.phase == 0) fresh.push_back(&b);
    if (b.phase == 2) L.fin_list.push_back(&b);
  }
  if (fresh.empty() && L.fin_list.empty()) {
    bool any = false;
    for (BItem& b : bt->items) any = any || (b.lane == li && b.phase == 1);
    if (!any) return 0;
  }
  // the lane's work follows what the caller enqueued on its stream before
  // this launch (input conversions, reuse of freed blocks)
  HIP_OK(hipStreamWaitEvent(s, bt->ev_caller, 0));
  int rc = start_members(bt, L, fresh);
  if (rc) return rc;
  for (BItem& b : bt->items)
    if (b.lane == li && b.phase == 1) L.act_list.push_back(&b);
  if (L.fin_list.empty() && L.act_list.empty()) return 0;
  clear_escalated(L, L.act_list);
  // (make_batch keeps this order: it is already sorted by size)
  std::stable_sort(L.act_list.begin(), L.act_list.end(),
                   [](const BItem* a, const BItem* b) { return a->p.Np > b->p.Np; });
  std::vector<BItem*> fac, fac32, nofac;
  split_factor(bt, L.act_list, fac, fac32, nofac);
  const bool fin = !L.fin_list.empty();
  if (fin) {
    // s2 follows everything enqueued on s so far, then runs beside the
    // Newton step; s joins it before the completion event
    HIP_OK(hipEventRecord(L.ev_fork, s));
    HIP_OK(hipStreamWaitEvent(L.s2, L.ev_fork, 0));
    rc = finalize_and_predict(bt, L, L.fin_list);
    if (rc) return rc;
    HIP_OK(hipEventRecord(L.ev_join, L.s2));
    for (BItem* b : L.fin_list) b->phase = 3;
  }
  if (!L.act_list.empty()) {
    std::vector<BItem*> later;
    for (BItem* b : L.act_list)
      if (b->iter > 0) later.push_back(b);
    const Batch ba = make_batch(L, 1, L.act_list);
    const Batch bl = make_batch(L, 9, later);
    const Batch bf = make_batch(L, 2, fac);
    const Batch bf32 = make_batch(L, 6, fac32);
    const Batch bn = make_batch(L, 8, nofac);
    newton_iter(ba, &bl, &bf, &bf32, &bn, prec, bt->refine, L.gather, s);
    HIP_OK(hipMemcpyAsync(L.hstep, L.gather, L.act_list.size() * 8, hipMemcpyDeviceToHost, s));
  }
  rc = check_launch();
  if (rc) return rc;
  if (fin) HIP_OK(hipStreamWaitEvent(s, L.ev_join, 0));
  HIP_OK(hipEventRecord(L.ev, s));
  L.launched = true;
  return 0;
}

// Results of lane `li`'s finished step: report its finalised members and
// advance the others.
static void lane_finish(pgp_batcher* bt, int li, int64_t* done_tags, int32_t max_done,
                        int32_t* n_done) {
  Lane& L = bt->lane[li];
  L.launched = false;
  report(bt, L.fin_list, L.hfin, done_tags, max_done, n_done);
  for (size_t i = 0; i < L.act_list.size(); ++i)
    if (advance(bt, L.act_list[i], L.hstep[i])) L.act_list[i]->phase = 2;
  L.fin_list.clear();
  L.act_list.clear();
  drop_done(bt);
}

// Members in flight from which a new member may take a lane past the
// second (PGP_LANE_SPLIT_MIN, default 512; 0: always): splitting the large
// members over more lanes raises the throughput of a full engine but slows
// the small members' lane while the engine is lightly loaded.
static int lane_split_min() {
  const char* e = getenv("PGP_LANE_SPLIT_MIN");
  return e ? std::max(0, atoi(e)) : 512;
}

// Size-class bounds on the padded N (PGP_LANE_NP: up to NLANES - 1
// ascending values separated by commas, default "1024,1536,2560"; "0" keeps
// every member in lane 0).  A member takes the lane of the number of bounds
// it exceeds.
static void lane_bounds(int* out) {
  for (int k = 0; k < NLANES - 1; ++k) out[k] = 0;
  const char* e = getenv("PGP_LANE_NP");
  const char* p = e ? e : "1024,1536,2560";
  for (int k = 0; k < NLANES - 1 && *p; ++k) {
    out[k] = atoi(p);
    while (*p && *p != ',') ++p;
    if (*p == ',') ++p;
  }
}

}  // namespace pgp

extern "C" {

int64_t pgp_batcher_table_size(int32_t capacity) {
  if (capacity <= 0) return -1;
  return batcher_table_bytes(capacity);
}

int pgp_batcher_create(int32_t capacity, const pgp_fit_options_t* opts, void* table,
                       int64_t table_bytes, void* stream, pgp_batcher_t** out) {
  if (capacity <= 0) return fail(-1, "capacity must be positive");
  if (!opts) return fail(-2, "options are NULL");
  if (opts->precision < 0 || opts->precision > 2) return fail(-2, "unknown precision");
  if (!table || table_bytes < batcher_table_bytes(capacity))
    return fail(-3, "batcher table too small");
  if (!out) return fail(-6, "out is NULL");
  pgp_batcher* b = new pgp_batcher();
  b->opts = *opts;
  b->refine = opts->refine >= 0 ? opts->refine : default_refine(opts->precision);
  b->cap = capacity;
  lane_bounds(b->lane_np);
  b->split_min = lane_split_min();
  bool ok = true;
  b->caller = (hipStream_t)stream;
  ok = hipEventCreateWithFlags(&b->ev_caller, hipEventDisableTiming) == hipSuccess;
  for (int li = 0; li < NLANES; ++li) {
    Lane& L = b->lane[li];
    char* t = (char*)table + li * lane_table_bytes(capacity);
    for (int i = 0; i < BATCH_ARRAYS; ++i) L.arr[i] = (Prob**)(t + (int64_t)i * capacity * 8);
    int64_t off = align256((int64_t)BATCH_ARRAYS * capacity * 8);
    L.gather = (double*)(t + off);
    off += align256((int64_t)capacity * 8 * 8);
    L.gather_fin = (double*)(t + off);
    off += align256((int64_t)capacity * 8 * 8);
    L.ticket = (uint32_t*)(t + off);
    L.ticket2 = L.ticket + 32;
    off += 256;
    L.ddesc = (Prob*)(t + off);
    // every lane runs on streams of its own (default priority); the caller's
    // stream only orders the lanes' work after the caller's
    ok = ok && hipStreamCreateWithFlags(&L.s, hipStreamNonBlocking) == hipSuccess;
    L.own_s = ok;
    ok = ok && hipMemsetAsync(L.ticket, 0, 256, (hipStream_t)stream) == hipSuccess &&
         hipHostMalloc((void**)&L.hstep, (size_t)capacity * 8, hipHostMallocDefault) ==
             hipSuccess &&
         hipHostMalloc((void**)&L.hfin, (size_t)capacity * 64, hipHostMallocDefault) ==
             hipSuccess &&
         hipHostMalloc((void**)&L.hdesc, (size_t)2 * capacity * sizeof(Prob),
                       hipHostMallocDefault) == hipSuccess &&
         hipEventCreateWithFlags(&L.ev, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&L.ev_fork, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&L.ev_join, hipEventDisableTiming) == hipSuccess &&
         hipStreamCreateWithFlags(&L.s2, hipStreamNonBlocking) == hipSuccess;
  }
  if (!ok || hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
    pgp_batcher_destroy(b);
    return fail(-100, "batcher set-up (table clear, pinned buffers, events, streams) failed");
  }
  *out = b;
  return 0;
}

int pgp_batcher_submit(pgp_batcher_t* bt, pgp_batch_item_t* item, int64_t tag) {
  if (!bt) return fail(-1, "batcher is NULL");
  if (!item) return fail(-2, "item is NULL");
  int rc = validate(&item->problem, &item->kernel, &item->lik);
  if (rc) return rc;
  const int prec = bt->opts.precision;
  const WsLayout L = ws_layout(item->problem.N, prec);
  if (!item->workspace || item->workspace_bytes < L.total)
    return fail(-2, "item workspace too small: %lld < %lld", (long long)item->workspace_bytes,
                (long long)L.total);
  if (item->Xtest) {
    if (item->M <= 0 || !item->mean || !item->var) return fail(-2, "predict outputs invalid");
    const PredLayout PL = pred_layout(L.Np, item->M, prec);
    if (!item->predict_workspace || item->predict_workspace_bytes < PL.total)
      return fail(-2, "item predict workspace too small");
  }
  BItem b;
  b.item = item;
  b.tag = tag;
  b.ws = (char*)item->workspace;
  b.p = make_prob(&item->problem, &item->kernel, &item->lik, prec, b.ws);
  b.phase = 0;
  b.iter = 0;
  b.lane = 0;
  b.done = false;
  b.step = INFINITY;
  b.gauss = item->lik.kind == PGP_LIK_GAUSSIAN;
  memset(&item->result, 0, sizeof item->result);
  bt->pending.push_back(b);
  return 0;
}

int32_t pgp_batcher_in_flight(pgp_batcher_t* bt) {
  return bt ? (int32_t)(bt->items.size() + bt->pending.size()) : 0;
}

int32_t pgp_batcher_launched(pgp_batcher_t* bt) {
  if (!bt) return 0;
  int32_t n = 0;
  for (int li = 0; li < NLANES; ++li) n += bt->lane[li].launched ? 1 : 0;
  return n;
}

int pgp_batcher_step(pgp_batcher_t* bt, int64_t* done_tags, int32_t max_done, int32_t* n_done) {
  if (!bt) return fail(-1, "batcher is NULL");
  if (!n_done) return fail(-4, "n_done is NULL");
  *n_done = 0;
  if (pgp_batcher_launched(bt)) return fail(-1, "pgp_batcher_step: an asynchronous launch is pending");
  // every member in one step, on lane 0's streams, after the caller's work
  Lane& L = bt->lane[0];
  hipStream_t s = L.s;
  const int prec = bt->opts.precision;
  admit(bt);
  if (bt->items.empty()) return 0;
  HIP_OK(hipEventRecord(bt->ev_caller, bt->caller));
  HIP_OK(hipStreamWaitEvent(s, bt->ev_caller, 0));
  std::vector<BItem*> fresh, active, fac, fac32, nofac;
  for (BItem& b : bt->items)
    if (b.phase == 0) fresh.push_back(&b);
  int rc = start_members(bt, L, fresh);
  if (rc) return rc;
  for (BItem& b : bt->items) active.push_back(&b);
  clear_escalated(L, active);
  split_factor(bt, active, fac, fac32, nofac);
  std::vector<BItem*> later;
  for (BItem* b : active)
    if (b->iter > 0) later.push_back(b);
  const Batch ba = make_batch(L, 1, active);
  const Batch bl = make_batch(L, 9, later);
  const Batch bf = make_batch(L, 2, fac);
  const Batch bf32 = make_batch(L, 6, fac32);
  const Batch bn = make_batch(L, 8, nofac);
  newton_iter(ba, &bl, &bf, &bf32, &bn, prec, bt->refine, L.gather, s);
  rc = check_launch();
  if (rc) return rc;
  HIP_OK(hipMemcpyAsync(L.hstep, L.gather, active.size() * 8, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  std::vector<BItem*> conv;
  for (size_t i = 0; i < active.size(); ++i)
    if (advance(bt, active[i], L.hstep[i])) conv.push_back(active[i]);
  if (conv.empty()) return 0;
  // s is idle (synchronised above), so s2 starts after the Newton step
  rc = finalize_and_predict(bt, L, conv);
  if (rc) return rc;
  HIP_OK(hipStreamSynchronize(L.s2));
  report(bt, conv, L.hfin, done_tags, max_done, n_done);
  drop_done(bt);
  return 0;
}

int pgp_batcher_launch(pgp_batcher_t* bt) {
  if (!bt) return fail(-1, "batcher is NULL");
  const auto t0 = std::chrono::steady_clock::now();
  admit(bt);
  if (bt->items.empty()) return 0;
  HIP_OK(hipEventRecord(bt->ev_caller, bt->caller));
  for (int li = 0; li < NLANES; ++li) {
    if (bt->lane[li].launched) continue;
    const int rc = lane_launch(bt, li);
    if (rc) return rc;
    if (bt->lane[li].launched) ++bt->lane_steps;
  }
  bt->host_enqueue_s +=
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

int pgp_batcher_stats(pgp_batcher_t* bt, double* host_enqueue_s, double* host_wait_s,
                      double* host_report_s, int64_t* lane_steps) {
  if (!bt) return fail(-1, "batcher is NULL");
  if (host_enqueue_s) *host_enqueue_s = bt->host_enqueue_s;
  if (host_wait_s) *host_wait_s = bt->host_wait_s;
  if (host_report_s) *host_report_s = bt->host_report_s;
  if (lane_steps) *lane_steps = bt->lane_steps;
  return 0;
}

int pgp_batcher_finish(pgp_batcher_t* bt, int64_t* done_tags, int32_t max_done,
                       int32_t* n_done) {
  if (!bt) return fail(-1, "batcher is NULL");
  if (!n_done) return fail(-4, "n_done is NULL");
  *n_done = 0;
  if (!pgp_batcher_launched(bt)) return 0;
  // wait until at least one launched lane has finished its step
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    bool any = false;
    for (int li = 0; li < NLANES; ++li) {
      Lane& L = bt->lane[li];
      if (!L.launched) continue;
      const hipError_t e = hipEventQuery(L.ev);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) return fail(-100, "lane %d step failed: %s", li, hipGetErrorString(e));
      if (!any) {
        const auto t1 = std::chrono::steady_clock::now();
        bt->host_wait_s += std::chrono::duration<double>(t1 - t0).count();
      }
      const auto t2 = std::chrono::steady_clock::now();
      lane_finish(bt, li, done_tags, max_done, n_done);
      bt->host_report_s +=
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t2).count();
      any = true;
    }
    if (any) return 0;
    std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

int32_t pgp_batcher_ready(pgp_batcher_t* bt) {
  if (!bt) return 0;
  int32_t n = 0;
  for (int li = 0; li < NLANES; ++li) {
    const Lane& L = bt->lane[li];
    // a failed step counts as ready: pgp_batcher_finish reports it
    if (L.launched && hipEventQuery(L.ev) != hipErrorNotReady) ++n;
  }
  return n;
}

int pgp_batcher_destroy(pgp_batcher_t* bt) {
  if (!bt) return 0;
  for (int li = 0; li < NLANES; ++li) {
    Lane& L = bt->lane[li];
    if (L.launched && L.ev) (void)hipEventSynchronize(L.ev);
    if (L.s2) {
      (void)hipStreamSynchronize(L.s2);
      (void)hipStreamDestroy(L.s2);
    }
    if (L.own_s && L.s) {
      (void)hipStreamSynchronize(L.s);
      (void)hipStreamDestroy(L.s);
    }
    for (hipEvent_t e : {L.ev, L.ev_fork, L.ev_join})
      if (e) (void)hipEventDestroy(e);
    if (L.hstep) (void)hipHostFree(L.hstep);
    if (L.hfin) (void)hipHostFree(L.hfin);
    if (L.hdesc) (void)hipHostFree(L.hdesc);
  }
  if (bt->ev_caller) (void)hipEventDestroy(bt->ev_caller);
  delete bt;
  return 0;
}

}  // extern "C"
