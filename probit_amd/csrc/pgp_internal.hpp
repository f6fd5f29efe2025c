// Internal declarations shared by the HIP translation units of libprobit_amd.
//
// Data model: one GP fit is a "problem".  Its device-side descriptor (Prob)
// lives at the start of the problem's caller-owned workspace and holds the
// kernel/likelihood parameters plus pointers into the same workspace:
//   K    : Np x Np FP64 Gram matrix, column-major, ld = Np (zero padding)
//   L    : Np x Np factor storage (double, or float in mixed modes), lower
//   dinv : Np/64 FP64 inverses of the 64 x 64 diagonal blocks of L
//   vec  : NVEC vectors of Np doubles (weight, latent, derivatives, ...)
// Batched kernels take a device array of descriptor pointers
// (grid.y or grid.z = batch member), so a single fit is a batch of one.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/probit_amd.h"

namespace pgp {

constexpr int TILE = 64;  // matrix tile / Cholesky panel width
constexpr int NTHREADS = 256;

// named vector slots in a problem's workspace
enum Vec {
  V_Z = 0,  // weight a (Newton / FPI iterate)
  V_F,      // latent f = K a
  V_G,      // d1 = d log p / d f
  V_W,      // W = -d2 log p / d f^2 (unclamped: the returned precision)
  V_S,      // sqrt(max(W, 0))
  V_R,      // Newton residual r = d1 - a
  V_T,      // solve right-hand side / scratch
  V_U,      // solve output
  V_DL,     // Newton step delta
  V_Q,      // refinement residual
  V_LL,     // per-point log-likelihood
  V_D3,     // third derivative (gradient path)
  V_X1,     // scratch
  V_X2,     // scratch
  V_X3,     // scratch
  V_X4,     // scratch
  V_SF,     // sqrt(max(W, 0)) at the last factorisation of B (the Newton solve's)
  NVEC
};

// reduction slots (doubles) per problem
enum Red {
  R_STEP2 = 0,  // ||delta||^2
  R_SUMLL,      // sum log p
  R_FA,         // f . a
  R_LOGDET,     // sum log L_ii (FP64)
  R_NEGW,       // count of W < 0
  NRED = 16
};

// What a workspace's factor storage holds (Prob::state), so that entry points
// acting on a fitted workspace can refuse one that holds something else.
enum ResidentState {
  RS_NONE = 0,     // nothing usable
  RS_LAPLACE = 1,  // B = I + W^1/2 K W^1/2 at a Laplace (or caller-given) state
  RS_VB = 2,       // B at W = 1/sigma^2 after a VB fit (objective_VB pieces valid)
  RS_VBMAP = 3     // sigma^2 I + K (after the VB map): not a predict state
};

struct Prob {
  int32_t N, Np, D, prec;
  int32_t kern_kind, periodic, lik_kind, J;
  int32_t sp_bounds, M, Mp, state;
  double variance, noise_std;
  double ls[PGP_MAX_DIM];           // lengthscales
  double inv_ls2[PGP_MAX_DIM];      // 1 / l^2
  double pi_over_per[PGP_MAX_DIM];  // pi / period (0 if not periodic)
  double per[PGP_MAX_DIM];          // periods (0 if not periodic)
  double cut[PGP_MAX_CUT];
  const double* X;
  const double* y;
  double* K;
  void* L;
  double* logdiag;
  double* diagval;
  double* dinv;
  double* vec;    // NVEC * Np
  double* red;    // NRED
  int32_t* info;  // [4]: [0] PGP_INFO bits, [1] count W < 0
  double* out_weight;
  double* out_prec;
  double* out_latent;
  // predict
  const double* Xt;  // M x D row-major
  void* S;           // Np x Mp factor-precision right-hand sides
  double* mupart;    // (Np/64) x Mp partial means
  double* pmean;
  double* pvar;
  // gradient / VB objective
  double* C;  // Np x Np FP64 (C = W^1/2 B^-1 W^1/2, or (sigma^2 I + K)^-1)
  // single-launch triangular solves: (epoch << 32) | blocks published
  uint64_t* progress;
  // symmetric products: per-tile partial sums, (Np/64) x (Np/64) x 64
  double* symv;
  // wavefront TRSM: one progress word per right-hand-side column tile (4)
  uint64_t* progress_ct;
  // predict with the FP64 variance correction: (Np/64) x Mp partial sums
  // (NULL: the variance from the column norms of L^-1 S)
  double* qpart;
  // merged column kernel of the Cholesky: 16 (Np/64) tile flags and a
  // workgroup count, zero between launches
  uint32_t* csync;
  // BF16 factors: BF16 copy of the rows below a super-step of its panel
  // (row r at (r - first row below) * panel16_pitch), which the super-step's
  // outer update reads; NULL for FP64 / FP32 factors
  unsigned short* panel16;
  int32_t panel16_pitch;
};
// words of Prob::csync for padded size Np
__host__ __device__ constexpr inline int64_t csync_words(int64_t Np) { return 16 * (Np / 64) + 16; }

__host__ __device__ inline double* vslot(const Prob& p, int v) {
  return p.vec + (size_t)v * (size_t)p.Np;
}

struct Batch {
  Prob* const* dprobs;  // device array of device pointers to descriptors
  int count;            // batch members in this launch
  int maxNp;            // max padded size among members
  const int* hnp = nullptr;  // host copy of each member's Np (NULL: all maxNp)
  const int* hm = nullptr;   // host copy of each member's test-point count M
  uint32_t* ticket = nullptr;  // device counter handing out workgroup order (zero between launches)
  int np(int m) const { return hnp ? hnp[m] : maxNp; }
  int mcount(int m, int dflt) const { return hm ? hm[m] : dflt; }
};

// Launch schedule of a batched kernel: the workgroups of one launch are
// listed member after member (pre[m] = first workgroup of member m), so a
// launch holds exactly the workgroups that have work, whatever the spread of
// member sizes.  Passed by value as a kernel argument.
constexpr int MAXB = 512;  // members per launch; larger batches are split
struct Sched {
  int32_t count;
  int32_t pre[MAXB + 1];
};

// Member of workgroup `bid` (the largest m with pre[m] <= bid) and the
// workgroup's index within that member.  Uniform: scalar loads only.
__device__ __forceinline__ int sched_member(const Sched& sc, int bid, int& local) {
  int lo = 0, hi = sc.count;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (sc.pre[mid] <= bid)
      lo = mid;
    else
      hi = mid;
  }
  local = bid - sc.pre[lo];
  return lo;
}

// Build the schedule(s) of a launch: units(m) = workgroups member m needs;
// go(dprobs, sched, total) launches one chunk of at most MAXB members.
template <class Units, class Go>
inline void sched_launch(const Batch& b, Units units, Go go) {
  for (int off = 0; off < b.count; off += MAXB) {
    const int n = b.count - off < MAXB ? b.count - off : MAXB;
    Sched sc;
    sc.count = n;
    int t = 0;
    for (int m = 0; m < n; ++m) {
      sc.pre[m] = t;
      t += units(off + m);
    }
    sc.pre[n] = t;
    if (t > 0) go(b.dprobs + off, sc, t);
  }
}

// Cross-workgroup publication inside one launch (wavefront solves, column
// steps of the factorisation): a workgroup only waits for workgroups that
// took a lower ticket (already running), and every wait is bounded.
constexpr int SPIN_LIMIT = 1 << 22;
uint32_t next_epoch();  // fresh epoch per launch (never 0)

// The same pointer in the global address space.  Loads through a generic
// pointer compile to FLAT instructions, which count against the LDS wait
// counter too: a kernel that prefetches from global memory while it reads
// LDS would then wait for the prefetch at its first LDS read.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gbl(T* p) {
  return (__attribute__((address_space(1))) T*)(uintptr_t)p;
}
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gbl(const T* p) {
  return (const __attribute__((address_space(1))) T*)(uintptr_t)p;
}

// Values handed between workgroups of one launch travel only through
// relaxed agent-scope accesses: the stores write through the XCD's L2 and
// the loads read past it, so a publication needs no L2 write-back or
// invalidation (the agent-scope fences, which act on the whole L2 of the
// XCD and slow every workgroup there).  A writer makes its coherent stores
// complete (publish_wait) before it stores the progress word; a reader
// issues its coherent loads after it has seen the word.
__device__ __forceinline__ double ld_coherent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coherent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The same for factor values in the storage type (float or double), through
// the global address space.
template <typename T>
__device__ __forceinline__ T ld_coh(const __attribute__((address_space(1))) T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_coh(T* p, T v) {
  __hip_atomic_store(gbl(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// This wave's outstanding global stores (and loads) complete; also a
// compiler barrier for memory.
__device__ __forceinline__ void publish_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Ticket of this workgroup in start order (thread 0 takes it; the last
// ticket resets the counter for the next launch on the stream).
__device__ __forceinline__ int take_ticket(uint32_t* ticket, int* lds_slot) {
  if (threadIdx.x == 0) {
    const uint32_t t = atomicAdd(ticket, 1u);
    if (t == gridDim.x - 1) atomicExch(ticket, 0u);
    *lds_slot = (int)t;
  }
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(*lds_slot);
}

// A launch whose ticket counter was not zero at its start has no blocks to
// hand out: its first n members (every member of the launch) are flagged
// PGP_INFO_STALLED, so their results are reported invalid, not silently
// skipped.  All threads of the workgroup call it.
__device__ __forceinline__ void flag_members_stalled(Prob* const* probs, int n) {
  for (int m = threadIdx.x; m < n; m += blockDim.x) atomicOr(probs[m]->info, PGP_INFO_STALLED);
}

// Wait (thread 0, bounded) until a member's progress word shows at least
// `need` published steps of launch `epoch`; flags PGP_INFO_STALLED if the
// bound runs out.  Returns the published count seen.
__device__ __forceinline__ int wait_progress(uint64_t* progress, uint32_t epoch, int need,
                                             int32_t* info) {
  int r = 0;
  for (int spins = 0;; ++spins) {
    const uint64_t pv = __hip_atomic_load(progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    r = ((uint32_t)(pv >> 32) == epoch) ? (int)(uint32_t)pv : 0;
    if (r >= need) break;
    if (spins >= SPIN_LIMIT) {
      atomicOr(info, PGP_INFO_STALLED);
      r = need;
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  // the published data is read with ld_coherent, issued after this point
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  return r;
}

// Opt-in launch timer: HIP events around the launches of one kernel class
// on the launching stream, with the algorithmic flops each launch performs.
//   TC_SYRK  outer (rank 64 G) trailing update of the Cholesky
//   TC_PANEL per-column Cholesky steps: diagonal factor, panel, column update
//   TC_TRSV  single-launch triangular solves
//   TC_GRAM  Gram assembly
//   TC_TRSM  multi-right-hand-side solves
//   TC_REFINE FP64 variance correction of predict (K U product)
//   TC_SUPER the merged column kernel (k_superstep)
enum TimerClass { TC_SYRK = 0, TC_PANEL, TC_TRSV, TC_GRAM, TC_TRSM, TC_REFINE, TC_SUPER,
                  TC_COUNT };
void timer_begin(int cls, hipStream_t s);
void timer_end(int cls, hipStream_t s, double algo_flops, double algo_bytes);
bool timer_on(int cls);

// gram.hip
void launch_gram_self(const Batch& b, hipStream_t s);
// (prec << 8) | vb; rs >= 0: also the partial sums of K x_rs (see
// launch_symv_reduce)
void launch_form_B(const Batch& b, int prec_vb, hipStream_t s, int rs = -1);
void launch_cross_prep(const Batch& b, int prec, int maxM, hipStream_t s);
void launch_pred_refine(const Batch& b, int prec, int maxM, hipStream_t s);
int launch_gram_raw(const pgp_kernel_t* k, int dtype, const double* X, int64_t N,
                    const double* Z, int64_t M, void* out, int64_t ld, hipStream_t s);

// lik.hip
void launch_lik_newton(const Batch& b, hipStream_t s);
void launch_lik_final(const Batch& b, int order3, hipStream_t s);
int launch_lik_raw(const pgp_likelihood_t* lik, const double* f, const double* y, int64_t N,
                   double* ll, double* d1, double* d2, hipStream_t s);
void launch_lik_grad_terms(const Batch& b, double* out, hipStream_t s);
int launch_probit_predictive(const pgp_likelihood_t* lik, const double* mean, const double* var,
                             int64_t M, double* out, hipStream_t s);
int launch_ordinal_metrics(const double* P, int J, const double* y, int64_t M, int32_t* y_pred,
                           double* plik, double* metrics, double* rows, hipStream_t s);

// chol.hip
void launch_potrf(const Batch& b, int prec, hipStream_t s);
int tune_class_of(int np);
pgp_tune_entry_t tune_get(int prec, int cls);
void tune_set(int prec, int cls, const pgp_tune_entry_t& v);
void launch_dinv(const Batch& b, int prec, hipStream_t s);
int launch_potrf_raw(int prec, void* A, int64_t N, int64_t ld, double* dinv, double* logdiag,
                     int32_t* info, hipStream_t s);

// Prior sampling workspace (padded size Np): K / its factor (Np x Np FP64),
// the diagonal-block inverses (Np x 64 FP64), the factor status word.
inline int64_t prior_ws_offset_dinv(int64_t Np) { return (Np * Np * 8 + 255) / 256 * 256; }
inline int64_t prior_ws_offset_info(int64_t Np) {
  return prior_ws_offset_dinv(Np) + (Np * TILE * 8 + 255) / 256 * 256;
}
inline int64_t prior_ws_bytes(int64_t Np) { return prior_ws_offset_info(Np) + 256; }
int launch_prior_sample(const pgp_kernel_t* kernel, const double* X, int64_t N, double jitter,
                        const double* z, int64_t S, int64_t ldz, double* f, int64_t ldf,
                        char* ws, hipStream_t s);
int launch_dinv_raw(int prec, const void* A, int64_t N, int64_t ld, double* dinv, hipStream_t s);

// solve.hip
void launch_trsv(const Batch& b, int prec, int trans, int in, int out, hipStream_t s);
void launch_trsm(const Batch& b, int prec, int maxMp, int lower_rhs, hipStream_t s);
void launch_trsm_t(const Batch& b, int prec, int maxMp, hipStream_t s);
int launch_trsv_raw_dinv(int prec, const void* L, int64_t N, int64_t ld, const double* dinv,
                         double* b, double* x, int trans, hipStream_t s);
int launch_trsm_raw_dinv(int prec, const void* L, int64_t N, int64_t ld, const double* dinv,
                         void* S, int64_t M, int64_t lds, int lower_rhs, hipStream_t s);

// vec.hip
enum VecOp {
  OP_STEP = 0,
  OP_IRCORR = 1,
  OP_ZUPD = 2,
  OP_ZERO_Z = 3,
  OP_OUT = 4,
  OP_VB_RHS = 5,
  OP_VB_STEP = 6,
  OP_VB_IRCORR = 7,
  OP_VB_PREC = 8,
  OP_GRAD_S2 = 9,    // X2 = -0.5 X1 .* D3
  OP_GRAD_SHAT = 10, // X3 = X2 - T
  // VB adjoint (pgp_vb_grad)
  OP_VBG_R = 11,     // X1 = Z - G
  OP_VBG_T = 12,     // T = S .* X2
  OP_VBG_Z = 13,     // X4 = S .* U
  OP_VBG_V = 14,     // X4 = (X2 - X3) / sigma^2
  OP_VBG_SHAT = 15,  // X3 = Z - G - sigma W .* X4, X1 = 0, X2 = sigma X4
  OP_SNAP = 16,      // SF = S: the W^1/2 of the factor B is about to be formed with
};
enum RedWhat { RED_STEP = 0, RED_FINAL = 1 };
void launch_gemv_sym(const Batch& b, int x_slot, int out_slot, int mode, hipStream_t s);
// the two halves of launch_gemv_sym's symmetric product (modes < 10)
void launch_symv_tiles(const Batch& b, int x_slot, hipStream_t s);
void launch_symv_reduce(const Batch& b, int x_slot, int out_slot, int mode, hipStream_t s);
void launch_vec_op(const Batch& b, int op, hipStream_t s);
void launch_copy_slot(const Batch& b, int src, int dst, hipStream_t s);
void launch_reduce(const Batch& b, int what, double* gather, hipStream_t s);
void launch_predict_final(const Batch& b, int prec, int maxM, hipStream_t s);
void launch_pred_cov(const Batch& b, int prec, int maxMp, double* cov, int64_t ld,
                     hipStream_t s);
// weight -> Z (if non-NULL), W -> W and S = sqrt(max(W, 0)) (if non-NULL); zero padding
void launch_load_state(const Batch& b, const double* weight, const double* W, hipStream_t s);

}  // namespace pgp
