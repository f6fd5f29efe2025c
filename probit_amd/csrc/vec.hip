// FP64 vector work of the Newton / fixed-point loops and the predict epilogue.
//
//  * symmetric GEMV y = K x: each lower 64x64 tile of K is read once (row
//    and column sums into per-tile slots, then a fixed-order reduction with
//    a fused epilogue: f = K a, S .* K r, refinement residuals), so HBM
//    reads half of K per product.
//    Replaces `K @ weight` (probit/implicit/Laplace.py:8,22,
//    probit/approximators.py:273,338) and the matvecs inside the reference's
//    dense Jacobian (probit/implicit/solvers.py:24).
//  * elementwise Newton updates and fixed-order per-problem reductions
//    (step norm for the jaxopt stopping rule, sum log p, f.a, log|B|).
//  * predict epilogue: mean_j = sum of per-tile partials, var_j = Kss -
//    ||V_:,j||^2 (probit/approximators.py:176-179).
#include "pgp_internal.hpp"

namespace pgp {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// mode 0: out = Kx
// mode 1: out = SF .* Kx              (SF: the W^1/2 of the factored B)
// mode 2: out = R - DL - S^2 .* Kx     (Laplace residual of (I + W K) d = r, x = DL)
// mode 3: out = X1 - sigma^2 x - Kx    (VB residual of (sigma^2 I + K) w = rhs, rhs in X1)
// (modes >= 10: the same with the matrix C in place of K)
__device__ __forceinline__ void gemv_epilogue(const Prob& P, int i, const double* x, int os,
                                              int mode, double sum) {
  double v;
  switch (mode) {
    case 0:
      v = sum;
      break;
    case 1:
      v = vslot(P, V_SF)[i] * sum;
      break;
    case 2: {
      const double si = vslot(P, V_S)[i];
      v = vslot(P, V_R)[i] - x[i] - si * si * sum;
      break;
    }
    default: {
      const double s2 = (i < P.N) ? P.noise_std * P.noise_std : 1.0;
      v = vslot(P, V_X1)[i] - s2 * x[i] - sum;
      break;
    }
  }
  vslot(P, os)[i] = v;
}

// Full-matrix product, one wave per output (used for the matrix C).
__global__ __launch_bounds__(NTHREADS) void k_gemv_sym(Prob* const* probs, Sched sc, int xs,
                                                       int os, int mode) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int i = bx * 4 + (threadIdx.x >> 6);
  if (i >= P.Np) return;
  const int lane = threadIdx.x & 63;
  const double* col = (mode >= 10 ? P.C : P.K) + (size_t)i * P.Np;
  mode %= 10;
  const double* x = vslot(P, xs);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int j = lane;
  for (; j + 192 < P.Np; j += 256) {
    a0 += col[j] * x[j];
    a1 += col[j + 64] * x[j + 64];
    a2 += col[j + 128] * x[j + 128];
    a3 += col[j + 192] * x[j + 192];
  }
  for (; j < P.Np; j += 64) a0 += col[j] * x[j];
  const double sum = wave_sum((a0 + a1) + (a2 + a3));
  if (lane == 0) gemv_epilogue(P, i, x, os, mode, sum);
}

// Symmetric product y = K x reading each lower 64x64 tile of K once: the
// tile (I, J) gives its row sums K_IJ x_J (towards y_I) and, off the
// diagonal, its column sums K_IJ^T x_I (towards y_J), each written to its
// own slot part[I][J] / part[J][I] (64 doubles).  k_symv_reduce then adds
// the nb slots of every row block in a fixed order, so the result does not
// depend on scheduling, and applies the epilogue.
__device__ __forceinline__ void lower_tile_v(int t, int& ti, int& tj) {
  int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((i + 1) * (i + 2) / 2 <= t) ++i;
  while (i * (i + 1) / 2 > t) --i;
  ti = i;
  tj = t - i * (i + 1) / 2;
}

__global__ __launch_bounds__(NTHREADS) void k_symv_tiles(Prob* const* probs, Sched sc, int xs) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int nb = P.Np / TILE;
  if (bx >= nb * (nb + 1) / 2) return;
  int ti, tj;
  lower_tile_v(bx, ti, tj);
  __shared__ double rpart[4][TILE];
  const size_t ld = (size_t)P.Np;
  const int r = threadIdx.x & 63, w = threadIdx.x >> 6;
  const double* x = vslot(P, xs);
  const double* Kt = P.K + (size_t)ti * TILE + (size_t)tj * TILE * ld;
  // lane r holds row r of the wave's 16 columns c = w + 4q (coalesced down
  // each column); the tile never goes through LDS
  double k[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) k[q] = Kt[(size_t)r + (size_t)(w + 4 * q) * ld];
  const double xi = x[ti * TILE + r];
  double rs = 0.0;
#pragma unroll
  for (int q = 0; q < 16; ++q) rs += k[q] * x[tj * TILE + w + 4 * q];
  rpart[w][r] = rs;
  double* part = P.symv;
  if (ti != tj) {
    // column sums K_IJ^T x_I of the wave's 16 columns: a transposing
    // butterfly halves the values per lane at each of lane bits 5..2, then
    // bits 1, 0 finish the sum; lane l ends with column q = (l >> 2) & 15
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = k[q] * xi;
#pragma unroll
    for (int h = 8, bit = 32; h >= 1; h >>= 1, bit >>= 1) {
      const bool up = (r & bit) != 0;
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
    if ((r & 3) == 0)
      part[((size_t)tj * nb + ti) * TILE + w + 4 * ((r >> 2) & 15)] = cs;
  }
  __syncthreads();
  if (w == 0)
    part[((size_t)ti * nb + tj) * TILE + r] = (rpart[0][r] + rpart[1][r]) + (rpart[2][r] + rpart[3][r]);
}

__global__ __launch_bounds__(NTHREADS) void k_symv_reduce(Prob* const* probs, Sched sc, int xs,
                                                          int os, int mode) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int i = bx * NTHREADS + threadIdx.x;
  if (i >= P.Np) return;
  const int nb = P.Np / TILE;
  const double* part = P.symv + (size_t)(i / TILE) * nb * TILE + (i % TILE);
  double a0 = 0.0, a1 = 0.0;
  int J = 0;
  for (; J + 1 < nb; J += 2) {
    a0 += part[(size_t)J * TILE];
    a1 += part[(size_t)(J + 1) * TILE];
  }
  if (J < nb) a0 += part[(size_t)J * TILE];
  gemv_epilogue(P, i, vslot(P, xs), os, mode, a0 + a1);
}

void launch_gemv_sym(const Batch& b, int x_slot, int out_slot, int mode, hipStream_t s) {
  if (mode >= 10) {
    sched_launch(
        b, [&](int m) { return (b.np(m) + 3) / 4; },
        [&](Prob* const* dp, const Sched& sc, int n) {
          hipLaunchKernelGGL(k_gemv_sym, dim3(n), dim3(NTHREADS), 0, s, dp, sc, x_slot, out_slot,
                             mode);
        });
    return;
  }
  launch_symv_tiles(b, x_slot, s);
  launch_symv_reduce(b, x_slot, out_slot, mode, s);
}

void launch_symv_tiles(const Batch& b, int x_slot, hipStream_t s) {
  sched_launch(
      b,
      [&](int m) {
        const int t = b.np(m) / TILE;
        return t * (t + 1) / 2;
      },
      [&](Prob* const* dp, const Sched& sc, int n) {
        hipLaunchKernelGGL(k_symv_tiles, dim3(n), dim3(NTHREADS), 0, s, dp, sc, x_slot);
      });
}

void launch_symv_reduce(const Batch& b, int x_slot, int out_slot, int mode, hipStream_t s) {
  sched_launch(
      b, [&](int m) { return (b.np(m) + NTHREADS - 1) / NTHREADS; },
      [&](Prob* const* dp, const Sched& sc, int n) {
        hipLaunchKernelGGL(k_symv_reduce, dim3(n), dim3(NTHREADS), 0, s, dp, sc, x_slot, out_slot,
                           mode);
      });
}



__global__ __launch_bounds__(NTHREADS) void k_vec(Prob* const* probs, Sched sc, int op) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int i = bx * NTHREADS + threadIdx.x;
  if (i >= P.Np) return;
  switch (op) {
    case OP_STEP:
      vslot(P, V_DL)[i] = vslot(P, V_R)[i] - vslot(P, V_SF)[i] * vslot(P, V_U)[i];
      break;
    case OP_IRCORR:
      vslot(P, V_DL)[i] += vslot(P, V_Q)[i] - vslot(P, V_SF)[i] * vslot(P, V_U)[i];
      break;
    case OP_SNAP:
      vslot(P, V_SF)[i] = vslot(P, V_S)[i];
      break;
    case OP_ZUPD:  // a step from a failed factor is not taken
      if (!(P.info[0] & PGP_INFO_NOT_PD)) vslot(P, V_Z)[i] += vslot(P, V_DL)[i];
      break;
    case OP_ZERO_Z:  // the zero weight and its latent f = K 0
      vslot(P, V_Z)[i] = 0.0;
      vslot(P, V_F)[i] = 0.0;
      break;
    case OP_OUT:
      if (i < P.N) {
        if (P.out_weight) P.out_weight[i] = vslot(P, V_Z)[i];
        if (P.out_prec) P.out_prec[i] = vslot(P, V_W)[i];
        if (P.out_latent) P.out_latent[i] = vslot(P, V_F)[i];
      }
      break;
    case OP_VB_RHS:
      vslot(P, V_X1)[i] = vslot(P, V_F)[i] + P.noise_std * vslot(P, V_G)[i];
      break;
    case OP_VB_STEP:
      vslot(P, V_DL)[i] = vslot(P, V_U)[i] - vslot(P, V_Z)[i];
      break;
    case OP_VB_IRCORR:
      vslot(P, V_U)[i] += vslot(P, V_X4)[i];
      break;
    case OP_GRAD_S2:
      vslot(P, V_X2)[i] = -0.5 * vslot(P, V_X1)[i] * vslot(P, V_D3)[i];
      break;
    case OP_GRAD_SHAT:
      vslot(P, V_X3)[i] = vslot(P, V_X2)[i] - vslot(P, V_T)[i];
      break;
    case OP_VBG_R:
      vslot(P, V_X1)[i] = vslot(P, V_Z)[i] - vslot(P, V_G)[i];
      break;
    case OP_VBG_T:
      vslot(P, V_T)[i] = vslot(P, V_S)[i] * vslot(P, V_X2)[i];
      break;
    case OP_VBG_Z:
      vslot(P, V_X4)[i] = vslot(P, V_S)[i] * vslot(P, V_U)[i];
      break;
    case OP_VBG_V:
      vslot(P, V_X4)[i] =
          (vslot(P, V_X2)[i] - vslot(P, V_X3)[i]) / (P.noise_std * P.noise_std);
      break;
    case OP_VBG_SHAT: {
      const double v = vslot(P, V_X4)[i];
      vslot(P, V_X3)[i] = vslot(P, V_Z)[i] - vslot(P, V_G)[i] - P.noise_std * vslot(P, V_W)[i] * v;
      vslot(P, V_X1)[i] = 0.0;
      vslot(P, V_X2)[i] = P.noise_std * v;
      break;
    }
    case OP_VB_PREC: {
      const double w = (i < P.N) ? 1.0 / (P.noise_std * P.noise_std) : 0.0;
      vslot(P, V_W)[i] = w;
      vslot(P, V_S)[i] = sqrt(w);
      break;
    }
    default:
      break;
  }
}

void launch_vec_op(const Batch& b, int op, hipStream_t s) {
  sched_launch(
      b, [&](int m) { return (b.np(m) + NTHREADS - 1) / NTHREADS; },
      [&](Prob* const* dp, const Sched& sc, int n) {
        hipLaunchKernelGGL(k_vec, dim3(n), dim3(NTHREADS), 0, s, dp, sc, op);
      });
}

__global__ __launch_bounds__(NTHREADS) void k_load_state(Prob* const* probs, Sched sc,
                                                         const double* weight, const double* W) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int i = bx * NTHREADS + threadIdx.x;
  if (i >= P.Np) return;
  if (weight) vslot(P, V_Z)[i] = (i < P.N) ? weight[i] : 0.0;
  if (W) {
    const double w = (i < P.N) ? W[i] : 0.0;
    vslot(P, V_W)[i] = w;
    vslot(P, V_S)[i] = sqrt(fmax(w, 0.0));
  }
}

void launch_load_state(const Batch& b, const double* weight, const double* W, hipStream_t s) {
  sched_launch(
      b, [&](int m) { return (b.np(m) + NTHREADS - 1) / NTHREADS; },
      [&](Prob* const* dp, const Sched& sc, int n) {
        hipLaunchKernelGGL(k_load_state, dim3(n), dim3(NTHREADS), 0, s, dp, sc, weight, W);
      });
}

// Copy a vector slot between problems' slots (same problem): dst = src.
__global__ __launch_bounds__(NTHREADS) void k_copy_slot(Prob* const* probs, Sched sc, int src,
                                                        int dst) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int i = bx * NTHREADS + threadIdx.x;
  if (i >= P.Np) return;
  vslot(P, dst)[i] = vslot(P, src)[i];
}

void launch_copy_slot(const Batch& b, int src, int dst, hipStream_t s) {
  sched_launch(
      b, [&](int m) { return (b.np(m) + NTHREADS - 1) / NTHREADS; },
      [&](Prob* const* dp, const Sched& sc, int n) {
        hipLaunchKernelGGL(k_copy_slot, dim3(n), dim3(NTHREADS), 0, s, dp, sc, src, dst);
      });
}

// ---- fixed-order block reduction ----
__device__ double block_sum(double v) {
  __shared__ double ws[NTHREADS / 64];
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  return (ws[0] + ws[1]) + (ws[2] + ws[3]);
}



__global__ __launch_bounds__(NTHREADS) void k_reduce(Prob* const* probs, int what,
                                                     double* gather) {
  const Prob& P = *probs[blockIdx.x];
  if (what == RED_STEP) {
    const double* d = vslot(P, V_DL);
    double a = 0.0;
    for (int i = threadIdx.x; i < P.Np; i += NTHREADS) a += d[i] * d[i];
    const double s = block_sum(a);
    if (threadIdx.x == 0) {
      P.red[R_STEP2] = s;
      if (gather) gather[blockIdx.x] = s;
    }
  } else {
    const double* ll = vslot(P, V_LL);
    const double* f = vslot(P, V_F);
    const double* z = vslot(P, V_Z);
    const double* w = vslot(P, V_W);
    double a = 0.0, b = 0.0, c = 0.0, n = 0.0;
    for (int i = threadIdx.x; i < P.Np; i += NTHREADS) {
      a += ll[i];
      b += f[i] * z[i];
      c += P.logdiag[i];
      n += (i < P.N && w[i] < 0.0) ? 1.0 : 0.0;
    }
    const double sa = block_sum(a);
    const double sb = block_sum(b);
    const double sc = block_sum(c);
    const double sn = block_sum(n);
    if (threadIdx.x == 0) {
      P.red[R_SUMLL] = sa;
      P.red[R_FA] = sb;
      P.red[R_LOGDET] = sc;
      P.red[R_NEGW] = sn;
      if (gather) {
        gather[8 * blockIdx.x + 0] = sa;
        gather[8 * blockIdx.x + 1] = sb;
        gather[8 * blockIdx.x + 2] = sc;
        gather[8 * blockIdx.x + 3] = sn;
        gather[8 * blockIdx.x + 4] = (double)P.info[0];
      }
    }
  }
}

void launch_reduce(const Batch& b, int what, double* gather, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce, dim3(b.count), dim3(NTHREADS), 0, s, b.dprobs, what, gather);
}

// ---- predict epilogue: one wave per test point ----
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_predict_final(Prob* const* probs, Sched sc) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int j = bx * 4 + (threadIdx.x >> 6);
  if (j >= P.M) return;
  const int lane = threadIdx.x & 63;
  const int nrt = P.Np / TILE;
  double q;
  if (P.qpart) {  // the FP64-corrected quadratic forms, summed over row tiles
    double qa = 0.0;
    for (int rt = lane; rt < nrt; rt += 64) qa += P.qpart[(size_t)rt * P.Mp + j];
    q = wave_sum(qa);
  } else {
    const T* S = (const T*)P.S + (size_t)j * P.Np;
    double q0 = 0.0, q1 = 0.0;
    int i = lane;
    for (; i + 64 < P.Np; i += 128) {
      const double v0 = (double)S[i], v1 = (double)S[i + 64];
      q0 += v0 * v0;
      q1 += v1 * v1;
    }
    for (; i < P.Np; i += 64) {
      const double v = (double)S[i];
      q0 += v * v;
    }
    q = wave_sum(q0 + q1);
  }
  double mu = 0.0;
  for (int rt = lane; rt < nrt; rt += 64) mu += P.mupart[(size_t)rt * P.Mp + j];
  mu = wave_sum(mu);
  if (lane == 0) {
    P.pmean[j] = mu;
    P.pvar[j] = P.variance - q;
  }
}

void launch_predict_final(const Batch& b, int prec, int maxM, hipStream_t s) {
  sched_launch(
      b, [&](int m) { return (b.mcount(m, maxM) + 3) / 4; },
      [&](Prob* const* dp, const Sched& sc, int n) {
        if (prec == PGP_PREC_FP64)
          hipLaunchKernelGGL(k_predict_final<double>, dim3(n), dim3(NTHREADS), 0, s, dp, sc);
        else
          hipLaunchKernelGGL(k_predict_final<float>, dim3(n), dim3(NTHREADS), 0, s, dp, sc);
      });
}

}  // namespace pgp
