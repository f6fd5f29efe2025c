// Gram-matrix assembly (mlkernels EQ / Matern12 / periodic / ARD semantics).
//
// Replaces `B.dense(prior(theta)(X))` (probit/implicit/Laplace.py:7,21,24,
// probit/implicit/VB.py:7,22, probit/approximators.py:174,272,337) and the
// cross-Gram `kernel(X, X_test)` (probit/approximators.py:173).
//
// Each workgroup owns one 64 x 64 output tile.  The two 64-row input slabs
// are staged in LDS, every thread evaluates one row x 16 columns, and the
// column-major output is written coalesced along rows.  Symmetric Gram
// matrices evaluate only the lower tiles and write the mirrored tile through
// an LDS transpose, so HBM sees each entry written once.
#include <math.h>

#include "mfma.hpp"
#include "pgp_internal.hpp"

namespace pgp {

struct KParams {
  int kind, periodic, D;
  double variance;
  double inv_ls2[PGP_MAX_DIM];
  double pi_over_per[PGP_MAX_DIM];
};

// Kernel parameters staged in LDS (broadcast reads, no per-lane copies).
struct KShared {
  double inv_ls2[PGP_MAX_DIM];
  double pi_over_per[PGP_MAX_DIM];
};

__device__ __forceinline__ void stage_params(KShared& ks, const double* inv_ls2,
                                             const double* pi_over_per) {
  if (threadIdx.x < PGP_MAX_DIM) {
    ks.inv_ls2[threadIdx.x] = inv_ls2[threadIdx.x];
    ks.pi_over_per[threadIdx.x] = pi_over_per[threadIdx.x];
  }
}

// k(x_i, z_j) from two LDS rows of length D.
template <typename TC>
__device__ __forceinline__ TC kernel_value(int kind, int periodic, int D, double variance,
                                           const KShared& ks, const TC* xi, const TC* xj) {
  TC r2 = 0;
  if (periodic) {
    for (int d = 0; d < D; ++d) {
      const TC s = sin((TC)ks.pi_over_per[d] * (xi[d] - xj[d]));
      r2 += (TC)4 * s * s * (TC)ks.inv_ls2[d];
    }
  } else {
    for (int d = 0; d < D; ++d) {
      const TC df = xi[d] - xj[d];
      r2 += df * df * (TC)ks.inv_ls2[d];
    }
  }
  const TC kap = (kind == PGP_KERNEL_EQ) ? exp((TC)-0.5 * r2) : exp(-sqrt(r2));
  return (TC)variance * kap;
}

// Decode a linear lower-triangular tile index t into (ti >= tj).
static inline int lower_tiles(int np) {
  const int t = np / TILE;
  return t * (t + 1) / 2;
}

__device__ __forceinline__ void lower_tile(int t, int& ti, int& tj) {
  int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((i + 1) * (i + 2) / 2 <= t) ++i;
  while (i * (i + 1) / 2 > t) --i;
  ti = i;
  tj = t - i * (i + 1) / 2;
}

// Load rows [r0, r0+64) of a row-major (rows x D) array into LDS (zero-padded).
template <typename TC>
__device__ __forceinline__ void load_rows(TC* dst, const double* src, int r0, int rows, int D) {
  for (int e = threadIdx.x; e < TILE * D; e += NTHREADS) {
    int r = e / D, d = e - r * D;
    int gr = r0 + r;
    dst[r * D + d] = (gr < rows) ? (TC)src[(size_t)gr * D + d] : (TC)0;
  }
}

// One 64x64 tile of a (possibly symmetric) Gram matrix.
//   out[i + j*ld] for i in [ti*64, ...), j in [tj*64, ...); entries outside
//   (nrows, ncols) are written as 0 if i < ldrows_pad and j < ncols_pad.
template <typename TC, typename TO>
__device__ void gram_tile(int kind, int periodic, int D, double variance, const double* inv_ls2,
                          const double* pi_over_per, const double* X, int nrows, const double* Z,
                          int ncols, TO* out, int64_t ld, int ti, int tj, int mirror,
                          int rows_pad, int cols_pad) {
  __shared__ TC xs[TILE * PGP_MAX_DIM];
  __shared__ TC zs[TILE * PGP_MAX_DIM];
  __shared__ TO st[TILE][TILE + 1];
  __shared__ KShared ks;
  stage_params(ks, inv_ls2, pi_over_per);
  load_rows(xs, X, ti * TILE, nrows, D);
  load_rows(zs, Z, tj * TILE, ncols, D);
  __syncthreads();
  const int il = threadIdx.x & 63;
  const int jg = threadIdx.x >> 6;
  const int i = ti * TILE + il;
  const TC* xi = &xs[il * D];
#pragma unroll 2
  for (int c = 0; c < 16; ++c) {
    const int jl = jg * 16 + c;
    const int j = tj * TILE + jl;
    TC v = 0;
    if (i < nrows && j < ncols)
      v = kernel_value<TC>(kind, periodic, D, variance, ks, xi, &zs[jl * D]);
    if (i < rows_pad && j < cols_pad) out[(size_t)i + (size_t)j * ld] = (TO)v;
    if (mirror) st[il][jl] = (TO)v;
  }
  if (mirror) {
    __syncthreads();
    // mirrored tile: rows tj*64 + jr, cols ti*64 + ic; value K[ti*64+ic][tj*64+jr]
    const int jr = threadIdx.x & 63;
    const int ig = threadIdx.x >> 6;
    for (int c = 0; c < 16; ++c) {
      const int ic = ig * 16 + c;
      const int row = tj * TILE + jr, col = ti * TILE + ic;
      if (row < rows_pad && col < cols_pad) out[(size_t)row + (size_t)col * ld] = st[ic][jr];
    }
  }
}

// ---- batched self Gram: K64 (Np x Np, full, zero padding) per problem ----
__global__ __launch_bounds__(NTHREADS) void k_gram_self(Prob* const* probs, Sched sc) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int T = P.Np / TILE;
  if (bx >= T * (T + 1) / 2) return;
  int ti, tj;
  lower_tile(bx, ti, tj);
  gram_tile<double, double>(P.kern_kind, P.periodic, P.D, P.variance, P.inv_ls2, P.pi_over_per,
                            P.X, P.N, P.X, P.N, P.K, P.Np, ti, tj, ti != tj, P.Np, P.Np);
}

void launch_gram_self(const Batch& b, hipStream_t s) {
  const bool on = timer_on(TC_GRAM);
  if (on) timer_begin(TC_GRAM, s);
  sched_launch(
      b, [&](int m) { return lower_tiles(b.np(m)); },
      [&](Prob* const* dp, const Sched& sc, int n) {
        hipLaunchKernelGGL(k_gram_self, dim3(n), dim3(NTHREADS), 0, s, dp, sc);
      });
  if (on) {
    // algorithmic bytes: the Np x Np FP64 matrix written once
    double bytes = 0.0;
    for (int m = 0; m < b.count; ++m) {
      const double np = b.np(m);
      bytes += np * np * 8.0;
    }
    timer_end(TC_GRAM, s, 0.0, bytes);
  }
}

// ---- B = I + S K S (Laplace) or sigma^2 I + K (VB) into the factor storage ----
// With rs >= 0 the same pass over K also yields the per-tile partial sums of
// the symmetric product K x, x = vector slot rs, in the slots k_symv_reduce
// adds (row sums of tile (I, J) in part[I][J], column sums in part[J][I];
// see k_symv_tiles), so the Newton step's K r costs no second read of K.
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_form_B(Prob* const* probs, Sched sc, int vb,
                                                     int rs) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int Tn = P.Np / TILE;
  if (bx >= Tn * (Tn + 1) / 2) return;
  int ti, tj;
  lower_tile(bx, ti, tj);
  const int il = threadIdx.x & 63;
  const int jg = threadIdx.x >> 6;
  const int i = ti * TILE + il;
  const int64_t ld = P.Np;
  T* L = (T*)P.L;
  const double* S = vslot(P, V_S);
  const double si = S[i];
  const double s2 = P.noise_std * P.noise_std;
  double k[16];
  for (int c = 0; c < 16; ++c) {
    const int j = tj * TILE + jg * 16 + c;
    const double kij = P.K[(size_t)i + (size_t)j * ld];
    k[c] = kij;
    double v;
    if (vb == 1) {  // sigma^2 I + K
      v = kij + ((i == j) ? ((i < P.N) ? s2 : 1.0) : 0.0);
    } else {        // s I + S K S: s = 1 (B), or sigma (vb == 2, the VB adjoint's B_sigma)
      v = si * kij * S[j] + ((i == j) ? ((vb == 2 && i < P.N) ? P.noise_std : 1.0) : 0.0);
    }
    L[(size_t)i + (size_t)j * ld] = (T)v;
  }
  if (rs < 0) return;
  __shared__ double rpart[4][TILE];
  const double* x = vslot(P, rs);
  double rsum = 0.0;
#pragma unroll
  for (int c = 0; c < 16; ++c) rsum += k[c] * x[tj * TILE + jg * 16 + c];
  rpart[jg][il] = rsum;
  double* part = P.symv;
  if (ti != tj) {
    // column sums of the wave's 16 columns by the transposing butterfly of
    // k_symv_tiles: lane l ends with column jg * 16 + ((l >> 2) & 15)
    const double xi = x[i];
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = k[q] * xi;
#pragma unroll
    for (int h = 8, bit = 32; h >= 1; h >>= 1, bit >>= 1) {
      const bool up = (il & bit) != 0;
#pragma unroll
      for (int q = 0; q < h; ++q) {
        const double give = up ? v[q] : v[q + h];
        const double keep = up ? v[q + h] : v[q];
        v[q] = keep + __shfl_xor(give, bit);
      }
    }
    double cs = v[0];
    cs += __shfl_xor(cs, 2);
    cs += __shfl_xor(cs, 1);
    if ((il & 3) == 0) part[((size_t)tj * Tn + ti) * TILE + jg * 16 + ((il >> 2) & 15)] = cs;
  }
  __syncthreads();
  if (jg == 0)
    part[((size_t)ti * Tn + tj) * TILE + il] =
        (rpart[0][il] + rpart[1][il]) + (rpart[2][il] + rpart[3][il]);
}

void launch_form_B(const Batch& b, int vb, hipStream_t s, int rs) {
  // all members of a batch share one storage precision (prec in vb's high bits)
  const int prec = vb >> 8;
  const int isvb = vb & 0xff;
  sched_launch(
      b, [&](int m) { return lower_tiles(b.np(m)); },
      [&](Prob* const* dp, const Sched& sc, int n) {
        if (prec == PGP_PREC_FP64)
          hipLaunchKernelGGL(k_form_B<double>, dim3(n), dim3(NTHREADS), 0, s, dp, sc, isvb, rs);
        else
          hipLaunchKernelGGL(k_form_B<float>, dim3(n), dim3(NTHREADS), 0, s, dp, sc, isvb, rs);
      });
}

// ---- cross Gram for predict: S = s .* Kfs (factor storage), partial means ----
// grid: x = (Np/64) * (Mp/64) tiles, y = batch member.  S (Np x Mp, ld Np)
// and mupart ((Np/64) x Mp doubles) live in the member's predict workspace.
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_cross_prep(Prob* const* probs, Sched sc) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int M = P.M;
  const int nrt = P.Np / TILE, nct = P.Mp / TILE;
  if (bx >= nrt * nct) return;
  const int ti = bx % nrt, tj = bx / nrt;
  __shared__ double xs[TILE * PGP_MAX_DIM];
  __shared__ double zs[TILE * PGP_MAX_DIM];
  __shared__ double part[4][16];
  __shared__ KShared ks;
  stage_params(ks, P.inv_ls2, P.pi_over_per);
  const int D = P.D;
  load_rows(xs, P.X, ti * TILE, P.N, D);
  load_rows(zs, P.Xt, tj * TILE, M, D);
  __syncthreads();
  const int il = threadIdx.x & 63;
  const int jg = threadIdx.x >> 6;
  const int i = ti * TILE + il;
  const double a = vslot(P, V_Z)[i];
  const double si = vslot(P, V_S)[i];
  T* S = (T*)P.S;
  const double* xi = &xs[il * D];
  double pv[16];
  for (int c = 0; c < 16; ++c) {
    const int jl = jg * 16 + c;
    const int j = tj * TILE + jl;
    double v = 0.0;
    if (i < P.N && j < M)
      v = kernel_value<double>(P.kern_kind, P.periodic, D, P.variance, ks, xi, &zs[jl * D]);
    S[(size_t)i + (size_t)j * P.Np] = (T)(si * v);
    pv[c] = v * a;
  }
  // sums of v * a over the 64 rows of the tile for the wave's 16 columns, by
  // a transposing butterfly (fixed order): lane l ends with column
  // jg * 16 + ((l >> 2) & 15)
#pragma unroll
  for (int h = 8, bit = 32; h >= 1; h >>= 1, bit >>= 1) {
    const bool up = (il & bit) != 0;
#pragma unroll
    for (int q = 0; q < h; ++q) {
      const double give = up ? pv[q] : pv[q + h];
      const double keep = up ? pv[q + h] : pv[q];
      pv[q] = keep + __shfl_xor(give, bit);
    }
  }
  double pa = pv[0];
  pa += __shfl_xor(pa, 2);
  pa += __shfl_xor(pa, 1);
  if ((il & 3) == 0) part[jg][(il >> 2) & 15] = pa;
  __syncthreads();
  if (threadIdx.x < TILE) {
    const int jg2 = threadIdx.x >> 4, c2 = threadIdx.x & 15;
    P.mupart[(size_t)ti * P.Mp + tj * TILE + threadIdx.x] = part[jg2][c2];
  }
}

void launch_cross_prep(const Batch& b, int prec, int maxM, hipStream_t s) {
  sched_launch(
      b,
      [&](int m) {
        const int mt = (b.mcount(m, maxM) + TILE - 1) / TILE;
        return (b.np(m) / TILE) * mt;
      },
      [&](Prob* const* dp, const Sched& sc, int n) {
        if (prec == PGP_PREC_FP64)
          hipLaunchKernelGGL(k_cross_prep<double>, dim3(n), dim3(NTHREADS), 0, s, dp, sc);
        else
          hipLaunchKernelGGL(k_cross_prep<float>, dim3(n), dim3(NTHREADS), 0, s, dp, sc);
      });
}

// ---- predictive covariance: cov = Kss - V^T V, V = L^-1 S after the TRSM ----
// One 64x64 lower tile (ti >= tj) of the M x M output per workgroup: an FP64
// MFMA product over the Np rows of V (FP32 values of V widen exactly), then
// Kss minus it, written to the tile and its mirror (the diagonal tile from
// its lower half), so the output is exactly symmetric.
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_pred_cov(Prob* const* probs, Sched sc, double* cov,
                                                       int64_t ld) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int Mt = P.Mp / TILE;
  if (bx >= Mt * (Mt + 1) / 2) return;
  int ti, tj;
  lower_tile(bx, ti, tj);
  constexpr int PA = TILE + 1;
  __shared__ double sa[TILE * PA], sb[TILE * PA];
  __shared__ double xs[TILE * PGP_MAX_DIM];
  __shared__ double zs[TILE * PGP_MAX_DIM];
  __shared__ KShared ks;
  const int D = P.D, M = P.M;
  stage_params(ks, P.inv_ls2, P.pi_over_per);
  load_rows(xs, P.Xt, ti * TILE, M, D);
  load_rows(zs, P.Xt, tj * TILE, M, D);
  const T* V = (const T*)P.S;
  const size_t np = (size_t)P.Np;
  const int kl = threadIdx.x & 63, cg = threadIdx.x >> 6;
  doublex4 acc[2][2];
  zero_acc<double>(acc);
  // k-major operands sa[k][m] = V(k0 + k, 64 ti + m), read coalesced along
  // k; the next chunk is fetched into registers while this one multiplies
  T ra[16], rb[16];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = cg + 4 * q;
      ra[q] = V[(size_t)(k0 + kl) + (size_t)(ti * TILE + c) * np];
      rb[q] = V[(size_t)(k0 + kl) + (size_t)(tj * TILE + c) * np];
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < P.Np; k0 += TILE) {
    __syncthreads();  // the previous chunk's products are done with LDS
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = cg + 4 * q;
      sa[kl * PA + c] = (double)ra[q];
      sb[kl * PA + c] = (double)rb[q];
    }
    __syncthreads();
    if (k0 + TILE < P.Np) fetch(k0 + TILE);
    mma64_kmajor<double>(sa, PA, sb, PA, acc);
  }
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = acc_row<double>(bm, r), col = acc_col(bn);
        const int i = ti * TILE + row, j = tj * TILE + col;
        if (i >= M || j >= M || (ti == tj && row < col)) continue;
        const double v =
            kernel_value<double>(P.kern_kind, P.periodic, D, P.variance, ks, &xs[row * D],
                                 &zs[col * D]) -
            acc[bm][bn][r];
        cov[(size_t)i + (size_t)j * ld] = v;
        cov[(size_t)j + (size_t)i * ld] = v;
      }
}

void launch_pred_cov(const Batch& b, int prec, int maxMp, double* cov, int64_t ld,
                     hipStream_t s) {
  const int mt = maxMp / TILE;
  sched_launch(
      b, [&](int) { return mt * (mt + 1) / 2; },
      [&](Prob* const* dp, const Sched& sc, int n) {
        if (prec == PGP_PREC_FP64)
          hipLaunchKernelGGL(k_pred_cov<double>, dim3(n), dim3(NTHREADS), 0, s, dp, sc, cov, ld);
        else
          hipLaunchKernelGGL(k_pred_cov<float>, dim3(n), dim3(NTHREADS), 0, s, dp, sc, cov, ld);
      });
}

// ---- FP64 correction of the predictive variance (factor stored in FP32) ----
// The variance needs q_j = s_j^T B^-1 s_j, B = I + W^1/2 K W^1/2, s_j the
// column j of S = W^1/2 Kfs.  The FP32 factor gives q_j = ||L^-1 s_j||^2 with
// the factor's rounding error; instead, with z_j = (L L^T)^-1 s_j from the
// two FP32 solves,
//   q_j = 2 s_j^T z_j - z_j^T B z_j = s_j^T B^-1 s_j - e^T B e,  e = z_j - B^-1 s_j,
// whose error is quadratic in the error of z_j.  With u = W^1/2 z,
// z^T B z = z^T z + u^T K u, and K is symmetric: u^T K u = sum_I u_I^T A_I with
// A_I = K_II u_I + 2 sum_{J < I} K_IJ u_J, so row tile I needs only the K
// tiles on and left of the diagonal.  One workgroup per 64 x 64 tile (I, J)
// of A, an FP64 MFMA product over those tiles of K (FP64), then per entry
// z (2 s - z) - u A with s recomputed in FP64 from the inputs, summed over the
// tile's 64 rows in a fixed order into qpart[I][j].  Z (the predict buffer
// after both solves) holds z in the factor's storage type.
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_pred_refine(Prob* const* probs, Sched sc) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int nrt = P.Np / TILE, nct = P.Mp / TILE;
  if (bx >= nrt * nct) return;
  const int ti = bx % nrt, tj = bx / nrt;
  constexpr int PA = TILE + 1;
  constexpr int KC = 32;  // k rows staged per step (half a tile: LDS and
                          // prefetch registers for more resident workgroups)
  __shared__ double sa[KC * PA], sb[KC * PA];
  __shared__ double red[2][TILE];
  __shared__ KShared ks;
  static_assert(TILE * PGP_MAX_DIM <= KC * PA, "input rows fit the operand buffers");
  const int D = P.D, M = P.M;
  const size_t np = (size_t)P.Np;
  const double* K = P.K;
  const T* Z = (const T*)P.S;
  const double* Ws = vslot(P, V_S);
  const int t = threadIdx.x;
  const int am = t & 63, ag = t >> 6;  // A: row m, k = ag + 4q
  const int bk = t & 31, bg = t >> 5;  // B: row k, n = bg + 8q
  doublex4 acc[2][2];
  zero_acc<double>(acc);
  // ra[q] = K(64 ti + am, k0 + ag + 4q); rb[q] = c U(k0 + bk, 64 tj + bg + 8q),
  // c = 2 left of the diagonal tile, 1 on it
  double ra[8], rb[8];
  const int kend = ti * TILE;  // the diagonal tile's first column
  auto fetch = [&](int k0) {
    const double w = (k0 < kend ? 2.0 : 1.0) * Ws[k0 + bk];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      ra[q] = K[(size_t)(ti * TILE + am) + (size_t)(k0 + ag + 4 * q) * np];
      rb[q] = w * (double)Z[(size_t)(k0 + bk) + (size_t)(tj * TILE + bg + 8 * q) * np];
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < kend + TILE; k0 += KC) {
    __syncthreads();  // the previous chunk's products are done with LDS
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      sa[(ag + 4 * q) * PA + am] = ra[q];  // A(m, k), k-major
      sb[bk * PA + bg + 8 * q] = rb[q];    // B(k, n), k-major
    }
    __syncthreads();
    if (k0 + KC < kend + TILE) fetch(k0 + KC);
    mma64_kmajor_k32<double>(sa, PA, sb, PA, acc);
  }
  __syncthreads();  // operands are dead: the input rows take their place
  double* xs = sa;
  double* zs = sb;
  stage_params(ks, P.inv_ls2, P.pi_over_per);
  load_rows(xs, P.X, ti * TILE, P.N, D);
  load_rows(zs, P.Xt, tj * TILE, M, D);
  __syncthreads();
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  double part[2] = {0.0, 0.0};
#pragma unroll
  for (int bn = 0; bn < 2; ++bn) {
    const int col = acc_col(bn);
    const int j = tj * TILE + col;
#pragma unroll
    for (int bm = 0; bm < 2; ++bm)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = acc_row<double>(bm, r);
        const int i = ti * TILE + row;
        if (i < P.N && j < M) {
          const double wi = Ws[i];
          const double z = (double)Z[(size_t)i + (size_t)j * np];
          const double sij = wi * kernel_value<double>(P.kern_kind, P.periodic, D, P.variance, ks,
                                                       &xs[row * D], &zs[col * D]);
          part[bn] += z * (2.0 * sij - z) - wi * z * acc[bm][bn][r];
        }
      }
  }
  // the 32 rows of this wave's half: lanes l, l + 16, l + 32, l + 48 hold
  // the same column
#pragma unroll
  for (int bn = 0; bn < 2; ++bn) {
    part[bn] += __shfl_xor(part[bn], 16);
    part[bn] += __shfl_xor(part[bn], 32);
  }
  if (l < 16) {
    red[w >> 1][32 * (w & 1) + l] = part[0];
    red[w >> 1][32 * (w & 1) + 16 + l] = part[1];
  }
  __syncthreads();
  if (threadIdx.x < TILE)
    P.qpart[(size_t)ti * P.Mp + tj * TILE + threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x];
}

void launch_pred_refine(const Batch& b, int prec, int maxM, hipStream_t s) {
  const bool on = timer_on(TC_REFINE);
  sched_launch(
      b,
      [&](int m) {
        const int mt = (b.mcount(m, maxM) + TILE - 1) / TILE;
        return (b.np(m) / TILE) * mt;
      },
      [&](Prob* const* dp, const Sched& sc, int n) {
        double fl = 0.0, by = 0.0;
        if (on) {
          // algorithmic work: the quadratic forms u^T K u over the lower
          // triangle of K, Np (Np + 64) Mp flops; bytes: the lower tiles of K
          // read once per column tile, Z read twice
          const int off = (int)(dp - b.dprobs);
          for (int m = off; m < off + sc.count; ++m) {
            const double np = b.np(m), mp = (b.mcount(m, maxM) + TILE - 1) / TILE * TILE;
            fl += np * (np + TILE) * mp;
            by += 0.5 * np * (np + TILE) * 8.0 * (mp / TILE) +
                  2.0 * np * mp * (prec == PGP_PREC_FP64 ? 8 : 4);
          }
          timer_begin(TC_REFINE, s);
        }
        if (prec == PGP_PREC_FP64)
          hipLaunchKernelGGL(k_pred_refine<double>, dim3(n), dim3(NTHREADS), 0, s, dp, sc);
        else
          hipLaunchKernelGGL(k_pred_refine<float>, dim3(n), dim3(NTHREADS), 0, s, dp, sc);
        if (on) timer_end(TC_REFINE, s, fl, by);
      });
}

// ---- raw Gram (kernel-level ABI) ----
template <typename TC, typename TO>
__global__ __launch_bounds__(NTHREADS) void k_gram_raw(KParams kp, const double* X, int N,
                                                       const double* Z, int M, TO* out,
                                                       int64_t ld, int sym) {
  int ti, tj;
  if (sym) {
    lower_tile(blockIdx.x, ti, tj);
  } else {
    const int nrt = (N + TILE - 1) / TILE;
    ti = blockIdx.x % nrt;
    tj = blockIdx.x / nrt;
  }
  gram_tile<TC, TO>(kp.kind, kp.periodic, kp.D, kp.variance, kp.inv_ls2, kp.pi_over_per, X, N, Z,
                    M, out, ld, ti, tj, sym && ti != tj, N, M);
}

KParams kparams_from_api(const pgp_kernel_t* k) {
  KParams kp;
  kp.kind = k->kind;
  kp.D = k->dim;
  kp.variance = k->variance;
  kp.periodic = 0;
  for (int d = 0; d < PGP_MAX_DIM; ++d) {
    kp.inv_ls2[d] = 0.0;
    kp.pi_over_per[d] = 0.0;
  }
  for (int d = 0; d < k->dim; ++d) {
    kp.inv_ls2[d] = 1.0 / (k->lengthscale[d] * k->lengthscale[d]);
    if (k->period[d] != 0.0) {
      kp.periodic = 1;
      kp.pi_over_per[d] = M_PI / k->period[d];
    }
  }
  return kp;
}

int launch_gram_raw(const pgp_kernel_t* k, int dtype, const double* X, int64_t N, const double* Z,
                    int64_t M, void* out, int64_t ld, hipStream_t s) {
  KParams kp = kparams_from_api(k);
  const int sym = (Z == X && M == N);
  const int nrt = (int)((N + TILE - 1) / TILE), nct = (int)((M + TILE - 1) / TILE);
  dim3 grid(sym ? nrt * (nrt + 1) / 2 : nrt * nct);
  if (dtype == 0)
    hipLaunchKernelGGL((k_gram_raw<double, double>), grid, dim3(NTHREADS), 0, s, kp, X, (int)N, Z,
                       (int)M, (double*)out, ld, sym);
  else
    hipLaunchKernelGGL((k_gram_raw<float, float>), grid, dim3(NTHREADS), 0, s, kp, X, (int)N, Z,
                       (int)M, (float*)out, ld, sym);
  return 0;
}

}  // namespace pgp
