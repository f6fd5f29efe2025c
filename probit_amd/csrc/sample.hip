// GP prior sampling: f = L_K z with L_K the Cholesky factor of K + jitter I
// (probit's data generators: examples/regression.py:52-58,
// examples/classification.py:227-234).  The Gram matrix and the factor come
// from the library's FP64 kernels; this file holds the two pieces around
// them: the jitter / padding of K and the lower-triangular product L z.
#include <hip/hip_runtime.h>

#include "mfma.hpp"
#include "pgp_internal.hpp"

namespace pgp {

// K (Np x Np, column-major, ld) after the Gram kernel wrote its N x N block:
// jitter on the diagonal, identity in the padding (rows / columns N..Np-1,
// so the padded factor is the identity there).  One workgroup per column;
// workgroup 0 also clears the factor's status word.
__global__ __launch_bounds__(NTHREADS) void k_prior_prepare(double* A, int64_t ld, int N, int Np,
                                                           double jitter, int32_t* info) {
  const int j = blockIdx.x;
  if (j == 0 && threadIdx.x == 0) *info = 0;
  double* col = A + (size_t)j * ld;
  if (j < N) {
    if (threadIdx.x == 0) col[j] += jitter;
    for (int i = N + threadIdx.x; i < Np; i += NTHREADS) col[i] = 0.0;
  } else {
    for (int i = threadIdx.x; i < Np; i += NTHREADS) col[i] = (i == j) ? 1.0 : 0.0;
  }
}

// f(64 I.., 64 c..) = sum_{J <= I} L(64 I.., 64 J..) z(64 J.., 64 c..) on
// FP64 MFMA, one workgroup per 64 x 64 output tile.  The diagonal block's
// upper triangle is masked (the factor storage holds only the lower
// triangle); rows of z past N and columns past S read as zero.
__global__ __launch_bounds__(NTHREADS) void k_trmm_lower(const double* L, int64_t ld, int N,
                                                        const double* z, int64_t ldz, int S,
                                                        double* f, int64_t ldf) {
  constexpr int PA = TILE + 1;
  __shared__ double As[TILE * PA], Bs[TILE * PA];
  const int I = blockIdx.x, c = blockIdx.y;
  const int m = threadIdx.x & 63, kb = threadIdx.x >> 6;
  const int64_t r0 = (int64_t)I * TILE, n0 = (int64_t)c * TILE;
  doublex4 acc[2][2];
  zero_acc<double>(acc);
  for (int J = 0; J <= I; ++J) {
    const int64_t k0 = (int64_t)J * TILE;
    __syncthreads();  // the previous product is done with LDS
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {
      const int k = kb + 4 * q;
      // As[k][m] = L(r0 + m, k0 + k), zero above the diagonal
      const double a = L[(size_t)(r0 + m) + (size_t)(k0 + k) * ld];
      As[k * PA + m] = (J == I && k > m) ? 0.0 : a;
      // Bs[k][n] = z(k0 + k, n0 + n): lane m reads row k0 + m of column n0 + k
      const bool ok = k0 + m < N && n0 + k < S;
      Bs[m * PA + k] = ok ? z[(size_t)(k0 + m) + (size_t)(n0 + k) * ldz] : 0.0;
    }
    __syncthreads();
    mma64_kmajor<double>(As, PA, Bs, PA, acc);
  }
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = r0 + acc_row<double>(bm, r), j = n0 + acc_col(bn);
        if (i < N && j < S) f[(size_t)i + (size_t)j * ldf] = acc[bm][bn][r];
      }
}

int launch_prior_sample(const pgp_kernel_t* kernel, const double* X, int64_t N, double jitter,
                        const double* z, int64_t S, int64_t ldz, double* f, int64_t ldf,
                        char* ws, hipStream_t s) {
  const int64_t Np = (N + TILE - 1) / TILE * TILE;
  double* A = (double*)ws;
  double* dinv = (double*)(ws + prior_ws_offset_dinv(Np));
  int32_t* info = (int32_t*)(ws + prior_ws_offset_info(Np));
  launch_gram_raw(kernel, 0, X, N, X, N, A, Np, s);
  hipLaunchKernelGGL(k_prior_prepare, dim3((unsigned)Np), dim3(NTHREADS), 0, s, A, Np, (int)N,
                     (int)Np, jitter, info);
  launch_potrf_raw(PGP_PREC_FP64, A, Np, Np, dinv, nullptr, info, s);
  hipLaunchKernelGGL(k_trmm_lower, dim3((unsigned)(Np / TILE), (unsigned)((S + TILE - 1) / TILE)),
                     dim3(NTHREADS), 0, s, (const double*)A, Np, (int)N, z, ldz, (int)S, f, ldf);
  return 0;
}

}  // namespace pgp
