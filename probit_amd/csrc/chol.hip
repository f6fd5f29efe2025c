// Blocked right-looking Cholesky A = L L^T (lower), replacing B.cholesky
// (probit/implicit/Laplace.py:24, probit/implicit/VB.py:10,25) and the LU
// solve of the Newton system (probit/implicit/solvers.py:24), whose matrix
// -(I + W K) is factored here in the symmetric form B = I + W^1/2 K W^1/2.
//
// Step k (column block c0 = 64k) is two launches over all batch members:
//   1. panel: every workgroup factors the 64x64 diagonal block in FP64
//      registers (redundantly) together with one 64-row block below it, so
//      the panel TRSM L21 = A21 L11^-T is fused into the factorisation;
//      workgroup 0 writes L11, log(L_ii) and the FP64 pivots.
//   2. trailing update A22 -= L21 L21^T on lower 64x64 tiles with MFMA:
//      FP64 (v_mfma_f64_16x16x4), FP32 (v_mfma_f32_16x16x4) or BF16
//      operands with FP32 accumulation (v_mfma_f32_16x16x32_bf16).
// After the last step one launch inverts every 64x64 diagonal block (FP64),
// which turns the diagonal solves of TRSV / TRSM into small products.
#include <math.h>

#include "mfma.hpp"
#include "pgp_internal.hpp"

namespace pgp {

__device__ __forceinline__ void lower_tile_c(int t, int& ti, int& tj) {
  int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((i + 1) * (i + 2) / 2 <= t) ++i;
  while (i * (i + 1) / 2 > t) --i;
  ti = i;
  tj = t - i * (i + 1) / 2;
}

// Factor the 128 x 64 panel [A11; A21] (A21 = rows r1..r1+63, absent if r1 < 0).
// Thread (tr = t & 15, tc = t >> 4) holds rows 8tr..8tr+7, cols 4tc..4tc+3.
template <typename T>
__device__ void panel_factor(T* A, int64_t ld, int64_t c0, int64_t r1, bool write_diag,
                             double* logdiag, double* diagval, int32_t* info) {
  __shared__ double colbuf[128];
  __shared__ double piv[64];
  const int t = threadIdx.x;
  const int tr = t & 15;
  const int tc = t >> 4;
  double a[8][4];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int R = 8 * tr + r;
    const bool valid = (R < 64) || (r1 >= 0);
    const int64_t grow = (R < 64) ? c0 + R : r1 + (R - 64);
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      const int64_t col = c0 + 4 * tc + cc;
      a[r][cc] = valid ? (double)A[(size_t)grow + (size_t)col * ld] : 0.0;
    }
  }
  // pivot (0, 0) is held by thread 0 as a[0][0]
  if (t == 0) piv[0] = a[0][0];
  __syncthreads();
  for (int j = 0; j < 64; ++j) {
    if (tc == (j >> 2)) {
      const double p = piv[j];
      const double d = sqrt(p);
      const double inv = 1.0 / d;
      const int jc = j & 3;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int R = 8 * tr + r;
        double cur = 0.0;
#pragma unroll
        for (int cc = 0; cc < 4; ++cc)
          if (cc == jc) cur = a[r][cc];
        double v;
        if (R < j)
          v = 0.0;
        else if (R == j)
          v = d;
        else
          v = cur * inv;
#pragma unroll
        for (int cc = 0; cc < 4; ++cc)
          if (cc == jc && R >= j) a[r][cc] = v;
        colbuf[R] = v;
        if (R == j && write_diag) {
          if (!(p > 0.0) && info) atomicOr(info, PGP_INFO_NOT_PD);
          if (logdiag) logdiag[c0 + j] = log(d);
          if (diagval) diagval[c0 + j] = d;
        }
      }
    }
    __syncthreads();
    double cb[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) cb[r] = colbuf[8 * tr + r];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      const int m = 4 * tc + cc;
      if (m > j) {
        const double cm = colbuf[m];
#pragma unroll
        for (int r = 0; r < 8; ++r) a[r][cc] -= cb[r] * cm;
      }
    }
    if (j + 1 < 64 && tr == ((j + 1) >> 3) && tc == ((j + 1) >> 2)) {
      const int rr = (j + 1) & 7, c1 = (j + 1) & 3;
#pragma unroll
      for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int cc = 0; cc < 4; ++cc)
          if (r == rr && cc == c1) piv[j + 1] = a[r][cc];
    }
    __syncthreads();
  }
  // write back
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int R = 8 * tr + r;
    if (R < 64) {
      if (!write_diag) continue;
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        const int m = 4 * tc + cc;
        if (m <= R) A[(size_t)(c0 + R) + (size_t)(c0 + m) * ld] = (T)a[r][cc];
      }
    } else if (r1 >= 0) {
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        const int m = 4 * tc + cc;
        A[(size_t)(r1 + R - 64) + (size_t)(c0 + m) * ld] = (T)a[r][cc];
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_potrf_panel(Prob* const* probs, int k) {
  const Prob& P = *probs[blockIdx.y];
  const int64_t c0 = (int64_t)k * TILE;
  if (c0 >= P.Np) return;
  const int nbk = (int)((P.Np - c0) / TILE);
  const int w = blockIdx.x;
  if (w >= (nbk > 1 ? nbk - 1 : 1)) return;
  const int64_t r1 = (nbk > 1) ? c0 + (int64_t)TILE * (w + 1) : -1;
  panel_factor<T>((T*)P.L, P.Np, c0, r1, w == 0, P.logdiag, P.diagval, P.info);
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_potrf_panel_raw(T* A, int64_t N, int64_t ld, int k,
                                                              double* logdiag, int32_t* info) {
  const int64_t c0 = (int64_t)k * TILE;
  const int nbk = (int)((N - c0) / TILE);
  const int w = blockIdx.x;
  const int64_t r1 = (nbk > 1) ? c0 + (int64_t)TILE * (w + 1) : -1;
  panel_factor<T>(A, ld, c0, r1, w == 0, logdiag, nullptr, info);
}

// ---- trailing update of one lower tile: C(ri.., rj..) -= P_i P_j^T ----
template <typename T, int BF16>
__device__ void syrk_tile(T* A, int64_t ld, int64_t c0, int64_t ri, int64_t rj) {
  typedef typename Acc<T>::type acc_t;
  acc_t acc[2][2];
  zero_acc<T>(acc);
  if constexpr (BF16) {
    __shared__ __attribute__((aligned(16))) unsigned short si[TILE * 72];
    __shared__ __attribute__((aligned(16))) unsigned short sj[TILE * 72];
    load_block_bf16_mmajor<T>(si, 72, A, ld, ri, c0);
    load_block_bf16_mmajor<T>(sj, 72, A, ld, rj, c0);
    __syncthreads();
    mma64_bf16(si, 72, sj, 72, *reinterpret_cast<floatx4(*)[2][2]>(&acc));
  } else {
    __shared__ T si[TILE * (TILE + 1)];
    __shared__ T sj[TILE * (TILE + 1)];
    load_block_kmajor<T, T>(si, TILE + 1, A, ld, ri, c0);
    load_block_kmajor<T, T>(sj, TILE + 1, A, ld, rj, c0);
    __syncthreads();
    mma64_kmajor<T>(si, TILE + 1, sj, TILE + 1, acc);
  }
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn) {
      const int64_t col = rj + acc_col(bn);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = ri + acc_row<T>(bm, r);
        T* p = &A[(size_t)row + (size_t)col * ld];
        *p = *p - acc[bm][bn][r];
      }
    }
}

template <typename T, int BF16>
__global__ __launch_bounds__(NTHREADS) void k_syrk(Prob* const* probs, int k) {
  const Prob& P = *probs[blockIdx.y];
  const int64_t c0 = (int64_t)k * TILE, c1 = c0 + TILE;
  if (c1 >= P.Np) return;
  const int Tk = (int)((P.Np - c1) / TILE);
  if ((int)blockIdx.x >= Tk * (Tk + 1) / 2) return;
  int ti, tj;
  lower_tile_c(blockIdx.x, ti, tj);
  syrk_tile<T, BF16>((T*)P.L, P.Np, c0, c1 + (int64_t)TILE * ti, c1 + (int64_t)TILE * tj);
}

template <typename T, int BF16>
__global__ __launch_bounds__(NTHREADS) void k_syrk_raw(T* A, int64_t N, int64_t ld, int k) {
  const int64_t c0 = (int64_t)k * TILE, c1 = c0 + TILE;
  int ti, tj;
  lower_tile_c(blockIdx.x, ti, tj);
  syrk_tile<T, BF16>(A, ld, c0, c1 + (int64_t)TILE * ti, c1 + (int64_t)TILE * tj);
}

// ---- inverse of every 64x64 diagonal block of the factor (FP64) ----
// DINV block k: column-major 64x64 at dinv + k * 4096.  Lane c of wave 0..
// computes column c of inv(L_kk) by forward substitution.
template <typename T>
__device__ void invert_diag_block(const T* A, int64_t ld, int64_t c0, double* out) {
  __shared__ double Ls[TILE][TILE + 1];  // Ls[col][row]
  const int m = threadIdx.x & 63, kb = threadIdx.x >> 6;
  for (int q = 0; q < 16; ++q) {
    const int col = kb + 4 * q;
    Ls[col][m] = (m >= col) ? (double)A[(size_t)(c0 + m) + (size_t)(c0 + col) * ld] : 0.0;
  }
  __syncthreads();
  // 4 waves x 16 columns each; lane = row index i of the solution column
  const int wv = threadIdx.x >> 6;
  const int i = threadIdx.x & 63;
  for (int cq = 0; cq < 16; ++cq) {
    const int c = wv * 16 + cq;
    // solve L x = e_c : x_j = (delta_jc - sum_{p<j} L_jp x_p) / L_jj
    double x = 0.0;  // lane i holds x_i
    double rhs = (i == c) ? 1.0 : 0.0;
    for (int j = c; j < 64; ++j) {
      double xj = 0.0;
      if (i == j) x = rhs / Ls[j][j];
      xj = __shfl(x, j);
      if (i > j) rhs -= Ls[j][i] * xj;
    }
    out[(size_t)i + (size_t)c * TILE] = (i >= c) ? x : 0.0;
  }
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_dinv(Prob* const* probs) {
  const Prob& P = *probs[blockIdx.y];
  const int nb = P.Np / TILE;
  if ((int)blockIdx.x >= nb) return;
  invert_diag_block<T>((const T*)P.L, P.Np, (int64_t)blockIdx.x * TILE,
                       P.dinv + (size_t)blockIdx.x * TILE * TILE);
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_dinv_raw(const T* A, int64_t ld, double* dinv) {
  invert_diag_block<T>(A, ld, (int64_t)blockIdx.x * TILE, dinv + (size_t)blockIdx.x * TILE * TILE);
}

// ---------------------------------------------------------------------------
template <typename T, int BF16>
static void potrf_steps(const Batch& b, hipStream_t s) {
  const int nb = b.maxNp / TILE;
  for (int k = 0; k < nb; ++k) {
    const int nbk = nb - k;
    dim3 gp(nbk > 1 ? nbk - 1 : 1, b.count);
    hipLaunchKernelGGL(k_potrf_panel<T>, gp, dim3(NTHREADS), 0, s, b.dprobs, k);
    const int Tk = nbk - 1;
    if (Tk > 0) {
      dim3 gs(Tk * (Tk + 1) / 2, b.count);
      hipLaunchKernelGGL((k_syrk<T, BF16>), gs, dim3(NTHREADS), 0, s, b.dprobs, k);
    }
  }
}

void launch_potrf(const Batch& b, int prec, hipStream_t s) {
  if (prec == PGP_PREC_FP64)
    potrf_steps<double, 0>(b, s);
  else if (prec == PGP_PREC_FP32)
    potrf_steps<float, 0>(b, s);
  else
    potrf_steps<float, 1>(b, s);
}

void launch_dinv(const Batch& b, int prec, hipStream_t s) {
  dim3 g(b.maxNp / TILE, b.count);
  if (prec == PGP_PREC_FP64)
    hipLaunchKernelGGL(k_dinv<double>, g, dim3(NTHREADS), 0, s, b.dprobs);
  else
    hipLaunchKernelGGL(k_dinv<float>, g, dim3(NTHREADS), 0, s, b.dprobs);
}

template <typename T, int BF16>
static void potrf_raw_steps(T* A, int64_t N, int64_t ld, double* logdiag, int32_t* info,
                            hipStream_t s) {
  const int nb = (int)(N / TILE);
  for (int k = 0; k < nb; ++k) {
    const int nbk = nb - k;
    hipLaunchKernelGGL(k_potrf_panel_raw<T>, dim3(nbk > 1 ? nbk - 1 : 1), dim3(NTHREADS), 0, s, A,
                       N, ld, k, logdiag, info);
    const int Tk = nbk - 1;
    if (Tk > 0)
      hipLaunchKernelGGL((k_syrk_raw<T, BF16>), dim3(Tk * (Tk + 1) / 2), dim3(NTHREADS), 0, s, A,
                         N, ld, k);
  }
}

int launch_potrf_raw(int prec, void* A, int64_t N, int64_t ld, double* logdiag, int32_t* info,
                     hipStream_t s) {
  if (prec == PGP_PREC_FP64)
    potrf_raw_steps<double, 0>((double*)A, N, ld, logdiag, info, s);
  else if (prec == PGP_PREC_FP32)
    potrf_raw_steps<float, 0>((float*)A, N, ld, logdiag, info, s);
  else
    potrf_raw_steps<float, 1>((float*)A, N, ld, logdiag, info, s);
  return 0;
}

int launch_dinv_raw(int prec, const void* A, int64_t N, int64_t ld, double* dinv, hipStream_t s) {
  if (prec == PGP_PREC_FP64)
    hipLaunchKernelGGL(k_dinv_raw<double>, dim3((unsigned)(N / TILE)), dim3(NTHREADS), 0, s,
                       (const double*)A, ld, dinv);
  else
    hipLaunchKernelGGL(k_dinv_raw<float>, dim3((unsigned)(N / TILE)), dim3(NTHREADS), 0, s,
                       (const float*)A, ld, dinv);
  return 0;
}

}  // namespace pgp
