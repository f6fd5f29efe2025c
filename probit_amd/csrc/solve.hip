// Triangular solves with the resident Cholesky factor.
//
// TRSV (one right-hand side per problem) serves the Newton step
// (I + W K)^-1 r = r - W^1/2 B^-1 W^1/2 K r (the reference's dense LU solve,
// probit/implicit/solvers.py:24) and B.cholesky_solve (probit/implicit/VB.py:11).
// TRSM (many right-hand sides) serves predict's solve against Kfs
// (probit/approximators.py:177) and the explicit inverses of
// probit/implicit/VB.py:26-27.
//
// Both are blocked by 64: one launch per block step.  The diagonal block is
// applied as a product with its precomputed FP64 inverse (DINV), so no step
// has a serial substitution chain; the off-diagonal update is a GEMV / MFMA
// GEMM.  Vectors are FP64 throughout; TRSM runs in the factor's storage type.
#include "mfma.hpp"
#include "pgp_internal.hpp"

namespace pgp {

// x = DINV_k b_k (trans = 0) or DINV_k^T b_k (trans = 1), 64 outputs in xs.
__device__ __forceinline__ void apply_dinv(const double* Dk, const double* bk_lds, double* xs,
                                           int trans) {
  __shared__ double part[4][TILE];
  const int i = threadIdx.x & 63, q = threadIdx.x >> 6;
  double acc = 0.0;
#pragma unroll 4
  for (int p = 16 * q; p < 16 * q + 16; ++p) {
    const double d = trans ? Dk[p + i * TILE] : Dk[i + p * TILE];
    acc += d * bk_lds[p];
  }
  part[q][i] = acc;
  __syncthreads();
  if (threadIdx.x < TILE) xs[i] = (part[0][i] + part[1][i]) + (part[2][i] + part[3][i]);
  __syncthreads();
}

template <typename T>
__device__ void trsv_fwd_step(const T* L, int64_t ld, int64_t Np, const double* dinv, int k,
                              double* b, double* x, int w) {
  __shared__ double bk[TILE], xk[TILE];
  __shared__ double part[4][TILE];
  const int64_t c0 = (int64_t)k * TILE;
  if (threadIdx.x < TILE) bk[threadIdx.x] = b[c0 + threadIdx.x];
  __syncthreads();
  apply_dinv(dinv + (size_t)k * TILE * TILE, bk, xk, 0);
  const int i = threadIdx.x & 63, q = threadIdx.x >> 6;
  if (w == 0) {
    if (threadIdx.x < TILE) x[c0 + i] = xk[i];
    return;
  }
  const int64_t R = c0 + (int64_t)TILE * w + i;
  double acc = 0.0;
#pragma unroll 4
  for (int p = 16 * q; p < 16 * q + 16; ++p)
    acc += (double)L[(size_t)R + (size_t)(c0 + p) * ld] * xk[p];
  part[q][i] = acc;
  __syncthreads();
  if (threadIdx.x < TILE) b[R] -= (part[0][i] + part[1][i]) + (part[2][i] + part[3][i]);
}

template <typename T>
__device__ void trsv_bwd_step(const T* L, int64_t ld, const double* dinv, int k, double* b,
                              double* x, int w) {
  __shared__ double bk[TILE], xk[TILE];
  const int64_t c0 = (int64_t)k * TILE;
  if (threadIdx.x < TILE) bk[threadIdx.x] = b[c0 + threadIdx.x];
  __syncthreads();
  apply_dinv(dinv + (size_t)k * TILE * TILE, bk, xk, 1);
  const int i = threadIdx.x & 63, q = threadIdx.x >> 6;
  if (w == 0) {
    if (threadIdx.x < TILE) x[c0 + i] = xk[i];
    return;
  }
  // columns m in [64 (w-1), 64 w): b_m -= sum_i L[c0 + i][m] x_i
  const double xi = xk[i];
  for (int cq = 0; cq < 16; ++cq) {
    const int64_t m = (int64_t)TILE * (w - 1) + 16 * q + cq;
    double v = (double)L[(size_t)(c0 + i) + (size_t)m * ld] * xi;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (i == 0) b[m] -= v;
  }
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_trsv(Prob* const* probs, int k, int trans, int in,
                                                   int out) {
  const Prob& P = *probs[blockIdx.y];
  const int nb = P.Np / TILE;
  if (k >= nb) return;
  const int w = blockIdx.x;
  double* b = vslot(P, in);
  double* x = vslot(P, out);
  if (!trans) {
    if (w >= nb - k) return;
    trsv_fwd_step<T>((const T*)P.L, P.Np, P.Np, P.dinv, k, b, x, w);
  } else {
    if (w > k) return;
    trsv_bwd_step<T>((const T*)P.L, P.Np, P.dinv, k, b, x, w);
  }
}

void launch_trsv(const Batch& bt, int prec, int trans, int in, int out, hipStream_t s) {
  const int nb = bt.maxNp / TILE;
  for (int step = 0; step < nb; ++step) {
    const int k = trans ? nb - 1 - step : step;
    dim3 g(trans ? k + 1 : nb - k, bt.count);
    if (prec == PGP_PREC_FP64)
      hipLaunchKernelGGL(k_trsv<double>, g, dim3(NTHREADS), 0, s, bt.dprobs, k, trans, in, out);
    else
      hipLaunchKernelGGL(k_trsv<float>, g, dim3(NTHREADS), 0, s, bt.dprobs, k, trans, in, out);
  }
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_trsv_raw(const T* L, int64_t N, int64_t ld,
                                                       const double* dinv, int k, int trans,
                                                       double* b, double* x) {
  if (!trans)
    trsv_fwd_step<T>(L, ld, N, dinv, k, b, x, blockIdx.x);
  else
    trsv_bwd_step<T>(L, ld, dinv, k, b, x, blockIdx.x);
}

int launch_trsv_raw_dinv(int prec, const void* L, int64_t N, int64_t ld, const double* dinv,
                         double* b, double* x, int trans, hipStream_t s) {
  const int nb = (int)(N / TILE);
  for (int step = 0; step < nb; ++step) {
    const int k = trans ? nb - 1 - step : step;
    dim3 g(trans ? k + 1 : nb - k);
    if (prec == PGP_PREC_FP64)
      hipLaunchKernelGGL(k_trsv_raw<double>, g, dim3(NTHREADS), 0, s, (const double*)L, N, ld,
                         dinv, k, trans, b, x);
    else
      hipLaunchKernelGGL(k_trsv_raw<float>, g, dim3(NTHREADS), 0, s, (const float*)L, N, ld, dinv,
                         k, trans, b, x);
  }
  return 0;
}

// ---- TRSM: S <- L^-1 S, S is Np x Mp (ld lds), one launch per block row ----
// WG (w, ct): X = DINV_k S_k[:, ct] (redundant per w); w == 0 stores X,
// w >= 1 updates S_w[:, ct] -= L_wk X.  lower_rhs: S is lower triangular
// (identity-like right-hand side): tiles with ct > k stay zero and are skipped.
template <typename T>
__device__ void trsm_step(const T* L, int64_t ld, const double* dinv, int k, T* S, int64_t lds,
                          int w, int ct, int lower_rhs) {
  __shared__ T As[TILE * (TILE + 1)];
  __shared__ T Bs[TILE * (TILE + 1)];
  typedef typename Acc<T>::type acc_t;
  const int64_t c0 = (int64_t)k * TILE;
  const int64_t n0 = (int64_t)ct * TILE;
  if (lower_rhs && ct > k) return;
  // A = DINV_k (k-major: As[kk*pitch + m] = Dinv(m, kk), Dinv column-major ld 64)
  load_block_kmajor<double, T>(As, TILE + 1, dinv + (size_t)k * TILE * TILE, TILE, 0, 0);
  // B[kk][n] = S(c0 + kk, n0 + n)
  {
    const int r = threadIdx.x & 63, cb = threadIdx.x >> 6;
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {
      const int c = cb + 4 * q;
      Bs[r * (TILE + 1) + c] = S[(size_t)(c0 + r) + (size_t)(n0 + c) * lds];
    }
  }
  __syncthreads();
  acc_t acc[2][2];
  zero_acc<T>(acc);
  mma64_kmajor<T>(As, TILE + 1, Bs, TILE + 1, acc);
  __syncthreads();
  if (w == 0) {
#pragma unroll
    for (int bm = 0; bm < 2; ++bm)
#pragma unroll
      for (int bn = 0; bn < 2; ++bn)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          S[(size_t)(c0 + acc_row<T>(bm, r)) + (size_t)(n0 + acc_col(bn)) * lds] = acc[bm][bn][r];
    return;
  }
  // X into Bs (k-major B operand: Bs[kk*pitch + n] = X(kk, n))
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r) Bs[acc_row<T>(bm, r) * (TILE + 1) + acc_col(bn)] = acc[bm][bn][r];
  const int64_t rw = c0 + (int64_t)TILE * w;
  load_block_kmajor<T, T>(As, TILE + 1, L, ld, rw, c0);
  __syncthreads();
  zero_acc<T>(acc);
  mma64_kmajor<T>(As, TILE + 1, Bs, TILE + 1, acc);
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        T* p = &S[(size_t)(rw + acc_row<T>(bm, r)) + (size_t)(n0 + acc_col(bn)) * lds];
        *p = *p - acc[bm][bn][r];
      }
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_trsm(Prob* const* probs, int k, int lower_rhs) {
  const Prob& P = *probs[blockIdx.z];
  const int nb = P.Np / TILE;
  if (k >= nb) return;
  const int w = blockIdx.x, ct = blockIdx.y;
  if (w >= nb - k || ct >= P.Mp / TILE) return;
  trsm_step<T>((const T*)P.L, P.Np, P.dinv, k, (T*)P.S, P.Np, w, ct, lower_rhs);
}

void launch_trsm(const Batch& bt, int prec, int maxMp, int lower_rhs, hipStream_t s) {
  const int nb = bt.maxNp / TILE;
  for (int k = 0; k < nb; ++k) {
    dim3 g(nb - k, maxMp / TILE, bt.count);
    if (prec == PGP_PREC_FP64)
      hipLaunchKernelGGL(k_trsm<double>, g, dim3(NTHREADS), 0, s, bt.dprobs, k, lower_rhs);
    else
      hipLaunchKernelGGL(k_trsm<float>, g, dim3(NTHREADS), 0, s, bt.dprobs, k, lower_rhs);
  }
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_trsm_raw(const T* L, int64_t ld, const double* dinv,
                                                       int k, T* S, int64_t lds, int lower_rhs) {
  trsm_step<T>(L, ld, dinv, k, S, lds, blockIdx.x, blockIdx.y, lower_rhs);
}

int launch_trsm_raw_dinv(int prec, const void* L, int64_t N, int64_t ld, const double* dinv,
                         void* S, int64_t M, int64_t lds, int lower_rhs, hipStream_t s) {
  const int nb = (int)(N / TILE);
  for (int k = 0; k < nb; ++k) {
    dim3 g(nb - k, (unsigned)(M / TILE));
    if (prec == PGP_PREC_FP64)
      hipLaunchKernelGGL(k_trsm_raw<double>, g, dim3(NTHREADS), 0, s, (const double*)L, ld, dinv,
                         k, (double*)S, lds, lower_rhs);
    else
      hipLaunchKernelGGL(k_trsm_raw<float>, g, dim3(NTHREADS), 0, s, (const float*)L, ld, dinv, k,
                         (float*)S, lds, lower_rhs);
  }
  return 0;
}

}  // namespace pgp
