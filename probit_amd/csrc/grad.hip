// Adjoint of the Laplace evidence and the VB traces.
//
// The reference differentiates objective_LA (probit/implicit/Laplace.py:12-30)
// with jax.value_and_grad through the implicit fixed-point layer
// (probit/implicit/solvers.py:28-64, probit/approximators.py:132-134).  At the
// fixed point that derivative has the closed form (Rasmussen & Williams
// Alg. 5.1, extended to likelihood parameters):
//   C = W^1/2 B^-1 W^1/2,  Sigma_ii = K_ii - (K C K)_ii,  s2 = -1/2 Sigma_ii d3_i,
//   shat = (I - C K) s2,
//   d/dtheta_K = sum_ij dK_ij/dtheta G_ij,  G = C/2 - a a^T/2 + (shat a^T + a shat^T)/2,
//   d/dtheta_l = sum_i [-L_p - Sigma_ii L_ffp / 2 + (K shat)_i L_fp].
// This file holds the O(N^3) pieces reused from the resident factor:
//   * k_sym_from_inv: C = S (Linv^T Linv) S (or Linv^T Linv for VB), full;
//   * k_grad_trace: tiles of sum_ij dK_ij/dtheta G_ij with dK recomputed from
//     X in registers, lower tiles only (off-diagonal tiles count twice),
//     per-tile partials reduced in a fixed order by k_sum_partials;
//   * k_vb_trace: tr(C) and sum_ij K_ij C_ij of objective_VB
//     (probit/implicit/VB.py:28-30).
#include <math.h>

#include "mfma.hpp"
#include "pgp_internal.hpp"

namespace pgp {

constexpr int GMAXD = 16;  // input dimensions supported by the gradient kernel
constexpr int NPAR = 1 + 2 * GMAXD;

__device__ __forceinline__ void lower_tile_g(int t, int& ti, int& tj) {
  int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((i + 1) * (i + 2) / 2 <= t) ++i;
  while (i * (i + 1) / 2 > t) --i;
  ti = i;
  tj = t - i * (i + 1) / 2;
}

// C(i0.., j0..) = scale_i * sum_k X(k, i) X(k, j) * scale_j, X column-major
// lower triangular (Linv), so the k range starts at max(i0, j0).
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_sym_from_inv(Prob* const* probs, int use_scale) {
  const Prob& P = *probs[blockIdx.y];
  const int nb = P.Np / TILE;
  if ((int)blockIdx.x >= nb * nb) return;
  const int ti = blockIdx.x % nb, tj = blockIdx.x / nb;
  const T* X = (const T*)P.S;  // Linv (Np x Np) lives in the gradient buffer
  const int64_t ld = P.Np;
  __shared__ T As[TILE * (TILE + 1)];
  __shared__ T Bs[TILE * (TILE + 1)];
  typedef typename Acc<T>::type acc_t;
  acc_t acc[2][2];
  zero_acc<T>(acc);
  const int kstart = ti > tj ? ti : tj;
  const int r = threadIdx.x & 63, cb = threadIdx.x >> 6;
  for (int kb = kstart; kb < nb; ++kb) {
    // As[kk][m] = X(kb*64 + kk, ti*64 + m); Bs[kk][n] = X(kb*64 + kk, tj*64 + n)
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {
      const int c = cb + 4 * q;
      As[r * (TILE + 1) + c] = X[(size_t)(kb * TILE + r) + (size_t)(ti * TILE + c) * ld];
      Bs[r * (TILE + 1) + c] = X[(size_t)(kb * TILE + r) + (size_t)(tj * TILE + c) * ld];
    }
    __syncthreads();
    mma64_kmajor<T>(As, TILE + 1, Bs, TILE + 1, acc);
    __syncthreads();
  }
  const double* S = vslot(P, V_S);
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int i = ti * TILE + acc_row<T>(bm, rr);
        const int j = tj * TILE + acc_col(bn);
        double v = (double)acc[bm][bn][rr];
        if (use_scale) v *= S[i] * S[j];
        if (i >= P.N || j >= P.N) v = 0.0;
        P.C[(size_t)i + (size_t)j * ld] = v;
      }
}

void launch_sym_from_inv(const Batch& b, int prec, int use_scale, hipStream_t s) {
  const int nb = b.maxNp / TILE;
  dim3 g(nb * nb, b.count);
  if (prec == PGP_PREC_FP64)
    hipLaunchKernelGGL(k_sym_from_inv<double>, g, dim3(NTHREADS), 0, s, b.dprobs, use_scale);
  else
    hipLaunchKernelGGL(k_sym_from_inv<float>, g, dim3(NTHREADS), 0, s, b.dprobs, use_scale);
}

// RHS for V = L^-1 (S .* K): G2(i, j) = S_i K_ij  (storage T); also the
// identity right-hand side for Linv (mode 1).
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_fill_rhs(Prob* const* probs, int mode) {
  const Prob& P = *probs[blockIdx.y];
  const int64_t n2 = (int64_t)P.Np * P.Np;
  const int64_t e = (int64_t)blockIdx.x * NTHREADS + threadIdx.x;
  if (e >= n2) return;
  const int i = (int)(e % P.Np), j = (int)(e / P.Np);
  T* out = (T*)P.S;
  double v;
  if (mode == 1)
    v = (i == j) ? 1.0 : 0.0;
  else
    v = vslot(P, V_S)[i] * P.K[e];
  out[e] = (T)v;
}

void launch_fill_rhs(const Batch& b, int prec, int mode, hipStream_t s) {
  const int64_t n2 = (int64_t)b.maxNp * b.maxNp;
  dim3 g((unsigned)((n2 + NTHREADS - 1) / NTHREADS), b.count);
  if (prec == PGP_PREC_FP64)
    hipLaunchKernelGGL(k_fill_rhs<double>, g, dim3(NTHREADS), 0, s, b.dprobs, mode);
  else
    hipLaunchKernelGGL(k_fill_rhs<float>, g, dim3(NTHREADS), 0, s, b.dprobs, mode);
}

// Sigma_jj = K_jj - ||V_:,j||^2 into X1 (one wave per column).
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_sigma_diag(Prob* const* probs) {
  const Prob& P = *probs[blockIdx.y];
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= P.Np) return;
  const int lane = threadIdx.x & 63;
  const T* V = (const T*)P.S + (size_t)j * P.Np;
  double q = 0.0;
  for (int i = lane; i < P.Np; i += 64) {
    const double v = (double)V[i];
    q += v * v;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) q += __shfl_xor(q, off);
  if (lane == 0) vslot(P, V_X1)[j] = (j < P.N) ? P.K[(size_t)j * P.Np + j] - q : 0.0;
}

void launch_sigma_diag(const Batch& b, int prec, hipStream_t s) {
  dim3 g((b.maxNp + 3) / 4, b.count);
  if (prec == PGP_PREC_FP64)
    hipLaunchKernelGGL(k_sigma_diag<double>, g, dim3(NTHREADS), 0, s, b.dprobs);
  else
    hipLaunchKernelGGL(k_sigma_diag<float>, g, dim3(NTHREADS), 0, s, b.dprobs);
}

// Fused trace: partials[tile][p] = w_tile * sum_{(i,j) in tile} dK_ij/dtheta_p G_ij
// with theta = (variance, lengthscale[D], period[D]) and G built from
// C (P.C), a (V_Z) and shat (V_X3).
__global__ __launch_bounds__(NTHREADS) void k_grad_trace(Prob* const* probs, double* partials) {
  const Prob& P = *probs[blockIdx.y];
  const int T = P.Np / TILE;
  const int ntiles = T * (T + 1) / 2;
  if ((int)blockIdx.x >= ntiles) return;
  int ti, tj;
  lower_tile_g(blockIdx.x, ti, tj);
  const int D = P.D;
  __shared__ double xs[TILE * GMAXD];
  __shared__ double zs[TILE * GMAXD];
  for (int e = threadIdx.x; e < TILE * D; e += NTHREADS) {
    const int r = e / D, d = e - r * D;
    const int gi = ti * TILE + r, gj = tj * TILE + r;
    xs[r * D + d] = gi < P.N ? P.X[(size_t)gi * D + d] : 0.0;
    zs[r * D + d] = gj < P.N ? P.X[(size_t)gj * D + d] : 0.0;
  }
  __syncthreads();
  const int il = threadIdx.x & 63, jg = threadIdx.x >> 6;
  const int i = ti * TILE + il;
  const double* a = vslot(P, V_Z);
  const double* sh = vslot(P, V_X3);
  const double ai = a[i], shi = sh[i];
  double acc[NPAR];
#pragma unroll
  for (int p = 0; p < NPAR; ++p) acc[p] = 0.0;
  const double wt = (ti == tj) ? 1.0 : 2.0;
  const bool eq = P.kern_kind == PGP_KERNEL_EQ;
  for (int c = 0; c < 16; ++c) {
    const int jl = jg * 16 + c;
    const int j = tj * TILE + jl;
    if (i >= P.N || j >= P.N) continue;
    const double G = wt * (0.5 * P.C[(size_t)i + (size_t)j * P.Np] - 0.5 * ai * a[j] +
                           0.5 * (shi * a[j] + ai * sh[j]));
    double q[GMAXD];
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < GMAXD; ++d) {
      q[d] = 0.0;
      if (d < D) {
        const double df = xs[il * D + d] - zs[jl * D + d];
        if (P.periodic) {
          const double sn = sin(P.pi_over_per[d] * df);
          q[d] = 4.0 * sn * sn;
        } else {
          q[d] = df * df;
        }
        r2 += q[d] * P.inv_ls2[d];
      }
    }
    double kap, dk_dr2;
    if (eq) {
      kap = exp(-0.5 * r2);
      dk_dr2 = -0.5 * P.variance * kap;
    } else {
      const double r = sqrt(r2);
      kap = exp(-r);
      dk_dr2 = (r > 0.0) ? -P.variance * kap / (2.0 * r) : 0.0;
    }
    acc[0] += G * kap;
    const double gd = G * dk_dr2;
#pragma unroll
    for (int d = 0; d < GMAXD; ++d) {
      if (d < D) {
        const double l = P.ls[d];
        acc[1 + d] += gd * (-2.0 * q[d] * P.inv_ls2[d] / l);
        if (P.periodic) {
          const double df = xs[il * D + d] - zs[jl * D + d];
          const double per = P.per[d];
          const double dq = -4.0 * sin(2.0 * P.pi_over_per[d] * df) * M_PI * df / (per * per);
          acc[1 + GMAXD + d] += gd * dq * P.inv_ls2[d];
        }
      }
    }
  }
  // fixed-order block reduction of each accumulator
  __shared__ double red[4];
  for (int p = 0; p < NPAR; ++p) {
    double v = acc[p];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
      partials[((size_t)blockIdx.y * ntiles + blockIdx.x) * NPAR + p] =
          (red[0] + red[1]) + (red[2] + red[3]);
  }
}

// out[b * NPAR + p] = sum over tiles (fixed order)
__global__ __launch_bounds__(NTHREADS) void k_sum_partials(Prob* const* probs,
                                                           const double* partials, double* out) {
  const Prob& P = *probs[blockIdx.x];
  const int T = P.Np / TILE;
  const int ntiles = T * (T + 1) / 2;
  __shared__ double red[NTHREADS];
  for (int p = 0; p < NPAR; ++p) {
    double v = 0.0;
    for (int t = threadIdx.x; t < ntiles; t += NTHREADS)
      v += partials[((size_t)blockIdx.x * ntiles + t) * NPAR + p];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int st = NTHREADS / 2; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)blockIdx.x * NPAR + p] = red[0];
    __syncthreads();
  }
}

int grad_npar() { return NPAR; }
int grad_maxd() { return GMAXD; }

void launch_grad_trace(const Batch& b, double* partials, double* out, hipStream_t s) {
  const int T = b.maxNp / TILE;
  dim3 g(T * (T + 1) / 2, b.count);
  hipLaunchKernelGGL(k_grad_trace, g, dim3(NTHREADS), 0, s, b.dprobs, partials);
  hipLaunchKernelGGL(k_sum_partials, dim3(b.count), dim3(NTHREADS), 0, s, b.dprobs, partials, out);
}

// VB traces: out[0] = tr(C), out[1] = sum_ij K_ij C_ij over the N x N block.
__global__ __launch_bounds__(NTHREADS) void k_vb_trace(Prob* const* probs, double* out) {
  const Prob& P = *probs[blockIdx.x];
  double tr = 0.0, kc = 0.0;
  const int64_t n2 = (int64_t)P.Np * P.Np;
  for (int64_t e = threadIdx.x; e < n2; e += NTHREADS) {
    const int i = (int)(e % P.Np), j = (int)(e / P.Np);
    if (i >= P.N || j >= P.N) continue;
    const double c = P.C[e];
    kc += P.K[e] * c;
    if (i == j) tr += c;
  }
  __shared__ double r1[NTHREADS], r2[NTHREADS];
  r1[threadIdx.x] = tr;
  r2[threadIdx.x] = kc;
  __syncthreads();
  for (int st = NTHREADS / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      r1[threadIdx.x] += r1[threadIdx.x + st];
      r2[threadIdx.x] += r2[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = r1[0];
    out[2 * blockIdx.x + 1] = r2[0];
  }
}

void launch_vb_trace(const Batch& b, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_vb_trace, dim3(b.count), dim3(NTHREADS), 0, s, b.dprobs, out);
}

// The VB adjoint's noise-std terms besides the likelihood partials:
// out[0] = tr(C), out[1] = v . (d1 - 2 sigma a) with v in X4.
__global__ __launch_bounds__(NTHREADS) void k_vb_sigma_terms(Prob* const* probs, double* out) {
  const Prob& P = *probs[blockIdx.x];
  const double* v = vslot(P, V_X4);
  const double* a = vslot(P, V_Z);
  const double* g = vslot(P, V_G);
  double tr = 0.0, dv = 0.0;
  for (int i = threadIdx.x; i < P.N; i += NTHREADS) {
    tr += P.C[(size_t)i + (size_t)i * P.Np];
    dv += v[i] * (g[i] - 2.0 * P.noise_std * a[i]);
  }
  __shared__ double r1[NTHREADS], r2[NTHREADS];
  r1[threadIdx.x] = tr;
  r2[threadIdx.x] = dv;
  __syncthreads();
  for (int st = NTHREADS / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      r1[threadIdx.x] += r1[threadIdx.x + st];
      r2[threadIdx.x] += r2[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = r1[0];
    out[2 * blockIdx.x + 1] = r2[0];
  }
}

void launch_vb_sigma_terms(const Batch& b, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_vb_sigma_terms, dim3(b.count), dim3(NTHREADS), 0, s, b.dprobs, out);
}

}  // namespace pgp
