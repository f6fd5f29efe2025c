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
#include <stdlib.h>

#include <algorithm>
#include <atomic>

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
__global__ __launch_bounds__(NTHREADS) void k_trsv(Prob* const* probs, Sched sc, int k, int trans,
                                                   int in, int out) {
  int w;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, w)];
  const int nb = P.Np / TILE;
  if (k >= nb) return;
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

// ---- single-launch wavefront TRSV --------------------------------------------
// One workgroup per 64-row block of every member, all in one launch.  In
// sweep order (forward: blocks 0, 1, ...; transposed: from the last block
// down), block i accumulates L_ij x_j for every block j before it as soon as
// x_j is published, then applies DINV_i and publishes x_i.  Publication is a
// per-member progress word ((epoch << 32) | blocks published) written with
// release semantics after x_i, read with acquire semantics; x travels through
// agent-scope (L2-coherent) loads and stores.  Workgroups take their block
// from a ticket counter in start order, so a workgroup only ever waits for
// workgroups that are already running: the wait always ends.  Tickets run
// step-major (`rows`: step s holds the members with more than s blocks, a
// prefix of the size-sorted batch), so the wave front moves through every
// member at once instead of member after member.  A wait is
// also bounded (SPIN_LIMIT polls); if the bound is ever reached the member
// is flagged PGP_INFO_STALLED and the sweep runs on, so every workgroup exits.

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_trsv_wave(Prob* const* probs, Sched rows,
                                                        uint32_t* ticket, uint32_t epoch,
                                                        int trans, int in, int out) {
  __shared__ double dk[TILE * (TILE + 1)];  // DINV_i, column-major, column pitch 65
  __shared__ double part[4][TILE];
  __shared__ double v[TILE];
  __shared__ int s_val;
  if (threadIdx.x == 0) {
    const uint32_t t = atomicAdd(ticket, 1u);
    if (t == gridDim.x - 1) atomicExch(ticket, 0u);  // every ticket taken: reset for the next launch
    s_val = (int)t;
  }
  __syncthreads();
  const int tk = __builtin_amdgcn_readfirstlane(s_val);
  if ((unsigned)tk >= gridDim.x) {  // counter not zero at launch: no block to take
    flag_members_stalled(probs, rows.pre[1] - rows.pre[0]);  // step 0 lists every member
    return;
  }
  int m;
  const int local = sched_member(rows, tk, m);  // step of this member's sweep
  const Prob& P = *probs[m];
  const int nb = P.Np / TILE;
  if (local >= nb) return;
  const int i = trans ? nb - 1 - local : local;
  const int64_t ld = P.Np;
  const T* L = (const T*)P.L;
  const double* b = vslot(P, in);
  double* x = vslot(P, out);
  const double* D = P.dinv + (size_t)i * TILE * TILE;
  for (int e = threadIdx.x; e < TILE * TILE; e += NTHREADS) dk[(e >> 6) * (TILE + 1) + (e & 63)] = D[e];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t ri = (int64_t)i * TILE;
  double pf = 0.0;      // forward: row l, columns 16w.. of each tile
  double pb[16];        // transposed: rows 16w + rr of block i, lane = row of block j
#pragma unroll
  for (int rr = 0; rr < 16; ++rr) pb[rr] = 0.0;
  int done = 0, ready = 0;
  while (done < local) {
    if (done >= ready) {
      __syncthreads();  // everyone has read the previous s_val
      if (threadIdx.x == 0) {
        // poll with relaxed loads; the published x is read with coherent
        // loads issued after the poll, so no cache maintenance is needed
        int r = 0;
        for (int spins = 0;; ++spins) {
          const uint64_t pv = __hip_atomic_load(P.progress, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
          r = ((uint32_t)(pv >> 32) == epoch) ? (int)(uint32_t)pv : 0;
          if (r > done) break;
          if (spins >= SPIN_LIMIT) {
            atomicOr(P.info, PGP_INFO_STALLED);
            r = local;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        s_val = r;
      }
      __syncthreads();
      ready = __builtin_amdgcn_readfirstlane(s_val);
    }
    const int lim = ready < local ? ready : local;
    // four published tiles at a time: their loads are in flight together
    for (; done + 4 <= lim; done += 4) {
      if (!trans) {
        double xv[4];
        T lv[4][16];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int64_t rj = (int64_t)(done + u) * TILE;
          xv[u] = ld_coherent(&x[rj + 16 * w + (l & 15)]);
#pragma unroll
          for (int c = 0; c < 16; ++c) lv[u][c] = L[(size_t)(ri + l) + (size_t)(rj + 16 * w + c) * ld];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int c = 0; c < 16; ++c) pf += (double)lv[u][c] * __shfl(xv[u], c);
      } else {
        double xl[4];
        T lv[4][16];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int64_t rj = (int64_t)(nb - 1 - (done + u)) * TILE;
          xl[u] = ld_coherent(&x[rj + l]);
#pragma unroll
          for (int rr = 0; rr < 16; ++rr)
            lv[u][rr] = L[(size_t)(rj + l) + (size_t)(ri + 16 * w + rr) * ld];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int rr = 0; rr < 16; ++rr) pb[rr] += (double)lv[u][rr] * xl[u];
      }
    }
    for (; done < lim; ++done) {
      const int j = trans ? nb - 1 - done : done;
      const int64_t rj = (int64_t)j * TILE;
      if (!trans) {
        const double xv = ld_coherent(&x[rj + 16 * w + (l & 15)]);
#pragma unroll
        for (int c = 0; c < 16; ++c) {
          const double xc = __shfl(xv, c);
          pf += (double)L[(size_t)(ri + l) + (size_t)(rj + 16 * w + c) * ld] * xc;
        }
      } else {
        const double xl = ld_coherent(&x[rj + l]);
#pragma unroll
        for (int rr = 0; rr < 16; ++rr)
          pb[rr] += (double)L[(size_t)(rj + l) + (size_t)(ri + 16 * w + rr) * ld] * xl;
      }
    }
  }
  // v = b_i - sum_j (...)
  if (!trans) {
    part[w][l] = pf;
  } else {
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) {
      double t = pb[rr];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
      if (l == 0) part[0][16 * w + rr] = t;
    }
  }
  __syncthreads();
  if (threadIdx.x < TILE) {
    const double acc = trans ? part[0][l] : (part[0][l] + part[1][l]) + (part[2][l] + part[3][l]);
    v[l] = b[ri + l] - acc;
  }
  __syncthreads();
  // x_i = DINV_i v (forward) or DINV_i^T v (transposed)
  double a = 0.0;
#pragma unroll
  for (int q = 16 * w; q < 16 * w + 16; ++q)
    a += (trans ? dk[q + l * (TILE + 1)] : dk[l + q * (TILE + 1)]) * v[q];
  part[w][l] = a;
  __syncthreads();
  if (threadIdx.x < TILE) {  // wave 0: store x_i, wait for the stores, then publish
    st_coherent(&x[ri + l], (part[0][l] + part[1][l]) + (part[2][l] + part[3][l]));
    publish_wait();
    if (threadIdx.x == 0)
      __hip_atomic_store(P.progress, ((uint64_t)epoch << 32) | (uint64_t)(local + 1),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

static bool wavefront_trsv() {
  static const bool v = [] {
    const char* e = getenv("PGP_TRSV_WAVEFRONT");
    return !(e && e[0] == '0');
  }();
  return v;
}

uint32_t next_epoch() {
  static std::atomic<uint32_t> g{0};
  uint32_t e = ++g;
  if (e == 0) e = ++g;
  return e;
}

void launch_trsv(const Batch& bt, int prec, int trans, int in, int out, hipStream_t s) {
  if (bt.ticket && wavefront_trsv() && bt.maxNp / TILE <= MAXB) {
    // step-major schedule: step s holds the members with more than s blocks
    // (a prefix: members are sorted by size), one workgroup each
    const int nb = bt.maxNp / TILE;
    Sched rows;
    rows.count = nb;
    int tot = 0;
    for (int st = 0; st < nb; ++st) {
      rows.pre[st] = tot;
      for (int m = 0; m < bt.count; ++m)
        if (bt.np(m) / TILE > st) ++tot;
    }
    rows.pre[nb] = tot;
    if (tot == 0) return;
    const bool on = timer_on(TC_TRSV);
    const double esz = prec == PGP_PREC_FP64 ? 8.0 : 4.0;
    // algorithmic work: Np^2 flops; bytes: the lower triangle of L and the
    // diagonal-block inverses, read once
    double fl = 0.0, by = 0.0;
    if (on) {
      for (int m = 0; m < bt.count; ++m) {
        const double np = bt.np(m);
        fl += np * np;
        by += 0.5 * np * (np + TILE) * esz + np * TILE * 8.0;
      }
      timer_begin(TC_TRSV, s);
    }
    const uint32_t ep = next_epoch();
    if (prec == PGP_PREC_FP64)
      hipLaunchKernelGGL(k_trsv_wave<double>, dim3(tot), dim3(NTHREADS), 0, s, bt.dprobs, rows,
                         bt.ticket, ep, trans, in, out);
    else
      hipLaunchKernelGGL(k_trsv_wave<float>, dim3(tot), dim3(NTHREADS), 0, s, bt.dprobs, rows,
                         bt.ticket, ep, trans, in, out);
    if (on) timer_end(TC_TRSV, s, fl, by);
    return;
  }
  const int nb = bt.maxNp / TILE;
  for (int step = 0; step < nb; ++step) {
    const int k = trans ? nb - 1 - step : step;
    sched_launch(
        bt,
        [&](int m) {
          const int nbm = bt.np(m) / TILE;
          return k < nbm ? (trans ? k + 1 : nbm - k) : 0;
        },
        [&](Prob* const* dp, const Sched& sc, int n) {
          if (prec == PGP_PREC_FP64)
            hipLaunchKernelGGL(k_trsv<double>, dim3(n), dim3(NTHREADS), 0, s, dp, sc, k, trans, in,
                               out);
          else
            hipLaunchKernelGGL(k_trsv<float>, dim3(n), dim3(NTHREADS), 0, s, dp, sc, k, trans, in,
                               out);
        });
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

// ---- TRSM: S <- L^-1 S, S is Np x Mp (ld lds) ----
// Two schedules.  For narrow right-hand sides (a few column tiles, as in
// serving) one launch per block row keeps every launch to two short tile
// products per workgroup; for wide ones the super-step schedule below reads
// and writes S G times less often.
//
// Per block row (lagged).
// Lagged schedule.  Launch k, workgroup (w, ct), row block i = k + w:
//   S_i[:, ct] -= L_{i,k-1} X_{k-1}[:, ct]   (k >= 1; X_{k-1} = S_{k-1}, final
//                                            since launch k - 1)
//   w == 0 then finishes block k: S_k[:, ct] = DINV_k S_k[:, ct].
// Every workgroup writes only its own rows and reads only rows finished in
// earlier launches, so workgroups of one launch never race, whatever order
// they run in.  Launch 0 holds only the w == 0 workgroups.
// lower_rhs: S is lower triangular (identity-like right-hand side): tiles
// with ct > k, and the updates of tile ct == k, are zero and are skipped.
template <typename T>
__device__ __forceinline__ void trsm_step(const T* L, int64_t ld, const double* dinv, int k, T* S, int64_t lds,
                          int w, int ct, int lower_rhs) {
  constexpr int PP = KP<T>::value;  // A operand: conflict-free MFMA reads
  constexpr int PB = TILE + 1;       // B operand: written transposed, conflict-free stores
  __shared__ T As[TILE * PP];
  __shared__ T Bs[TILE * PB];
  typedef typename Acc<T>::type acc_t;
  if (lower_rhs && (ct > k || (ct == k && w > 0))) return;
  const int64_t r0 = (int64_t)(k + w) * TILE;
  const int64_t n0 = (int64_t)ct * TILE;
  acc_t acc[2][2];
  zero_acc<T>(acc);
  if (k > 0) {
    const int64_t cp = (int64_t)(k - 1) * TILE;
    // A = L(r0.., cp..) (k-major), B[kk][n] = S(cp + kk, n0 + n)
    load_block_kmajor<T, T>(As, PP, L, ld, r0, cp);
    {
      const int r = threadIdx.x & 63, cb = threadIdx.x >> 6;
#pragma unroll 4
      for (int q = 0; q < 16; ++q) {
        const int c = cb + 4 * q;
        Bs[r * PB + c] = S[(size_t)(cp + r) + (size_t)(n0 + c) * lds];
      }
    }
    __syncthreads();
    mma64_kmajor<T>(As, PP, Bs, PB, acc);
    __syncthreads();
  }
  if (w > 0) {
#pragma unroll
    for (int bm = 0; bm < 2; ++bm)
#pragma unroll
      for (int bn = 0; bn < 2; ++bn)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          T* p = &S[(size_t)(r0 + acc_row<T>(bm, r)) + (size_t)(n0 + acc_col(bn)) * lds];
          *p = *p - acc[bm][bn][r];
        }
    return;
  }
  // block k: S_k - update (rounded to T) as the k-major B operand, DINV_k as A
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = acc_row<T>(bm, r), col = acc_col(bn);
        Bs[row * PB + col] =
            S[(size_t)(r0 + row) + (size_t)(n0 + col) * lds] - acc[bm][bn][r];
      }
  load_block_kmajor<double, T>(As, PP, dinv + (size_t)k * TILE * TILE, TILE, 0, 0);
  __syncthreads();
  zero_acc<T>(acc);
  mma64_kmajor<T>(As, PP, Bs, PB, acc);
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        S[(size_t)(r0 + acc_row<T>(bm, r)) + (size_t)(n0 + acc_col(bn)) * lds] = acc[bm][bn][r];
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_trsm_lagged(Prob* const* probs, Sched sc, int k,
                                                          int lower_rhs) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int nb = P.Np / TILE;
  if (k >= nb) return;
  const int rows = k == 0 ? 1 : nb - k;  // row blocks of this launch
  const int w = bx % rows, ct = bx / rows;
  if (ct >= P.Mp / TILE) return;
  trsm_step<T>((const T*)P.L, P.Np, P.dinv, k, (T*)P.S, P.Np, w, ct, lower_rhs);
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_trsm_lagged_raw(const T* L, int64_t ld,
                                                              const double* dinv, int k, T* S,
                                                              int64_t lds, int lower_rhs) {
  trsm_step<T>(L, ld, dinv, k, S, lds, blockIdx.x, blockIdx.y, lower_rhs);
}

// Wavefront TRSM for narrow right-hand sides (at most WAVE_NCT column
// tiles), all members in one launch: one workgroup per (block row i,
// member, column tile ct), numbered row-major so that the wave front moves
// through every member at once.  Workgroup (i, ct) accumulates L_ij X_j[:, ct]
// for every block j < i as soon as (j, ct) publishes, then forms
// X_i = DINV_i (S_i - acc) and publishes it: X_i through agent-scope stores,
// a wait for their completion, then the member's progress word of column
// tile ct.  Start-order tickets and bounded waits as in k_trsv_wave.
constexpr int WAVE_NCT = 4;

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_trsm_rows(Prob* const* probs, Sched rows, int nct,
                                                        uint32_t* ticket, uint32_t epoch) {
  constexpr int PP = KP<T>::value;  // A operand: conflict-free MFMA reads
  constexpr int PB = TILE + 1;       // B operand: written transposed, conflict-free stores
  __shared__ T As[TILE * PP];
  __shared__ T Bs[TILE * PB];
  __shared__ int s_val;
  typedef typename Acc<T>::type acc_t;
  const int tk = take_ticket(ticket, &s_val);
  if ((unsigned)tk >= gridDim.x) {  // counter not zero at launch: no block to take
    flag_members_stalled(probs, (rows.pre[1] - rows.pre[0]) / nct);  // row 0: every member
    return;
  }
  int local;
  const int i = sched_member(rows, tk, local);
  const int m = local / nct, ct = local - m * nct;
  const Prob& P = *probs[m];
  if (ct >= P.Mp / TILE || i >= P.Np / TILE) return;
  uint64_t* prog = P.progress_ct + ct;
  const auto L = gbl((const T*)P.L);
  T* S = (T*)P.S;
  const int64_t ld = P.Np;
  const int64_t ri = (int64_t)i * TILE, n0 = (int64_t)ct * TILE;
  const int r = threadIdx.x & 63, cb = threadIdx.x >> 6;
  // Operands that exist before the launch are fetched into registers ahead
  // of the waits: DINV_i and S_i at once, the next L_ij while X_j is awaited.
  T dv[16], la[16], sv[2][2][4];
  {
    const auto D = gbl(P.dinv + (size_t)i * TILE * TILE);
    const auto Sg = gbl((const T*)P.S);
#pragma unroll
    for (int q = 0; q < 16; ++q) dv[q] = (T)D[r + (size_t)(cb + 4 * q) * TILE];
#pragma unroll
    for (int bm = 0; bm < 2; ++bm)
#pragma unroll
      for (int bn = 0; bn < 2; ++bn)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          sv[bm][bn][q] =
              Sg[(size_t)(ri + acc_row<T>(bm, q)) + (size_t)(n0 + acc_col(bn)) * ld];
  }
  auto fetch_l = [&](int j) {
#pragma unroll
    for (int q = 0; q < 16; ++q)
      la[q] = L[(size_t)(ri + r) + (size_t)((int64_t)j * TILE + cb + 4 * q) * ld];
  };
  if (i > 0) fetch_l(0);
  acc_t acc[2][2];
  zero_acc<T>(acc);
  int done = 0, ready = 0;
  while (done < i) {
    if (done >= ready) {
      __syncthreads();  // everyone has read the previous s_val
      if (threadIdx.x == 0) s_val = wait_progress(prog, epoch, done + 1, P.info);
      __syncthreads();
      ready = min(__builtin_amdgcn_readfirstlane(s_val), i);
    }
    for (; done < ready; ++done) {
      const int64_t rj = (int64_t)done * TILE;
      __syncthreads();  // the previous product is done with LDS
#pragma unroll
      for (int q = 0; q < 16; ++q) As[(cb + 4 * q) * PP + r] = la[q];
      // B[kk][n] = X_j(kk, n), published in this launch: coherent loads
#pragma unroll 4
      for (int q = 0; q < 16; ++q) {
        const int c = cb + 4 * q;
        Bs[r * PB + c] = __hip_atomic_load(&S[(size_t)(rj + r) + (size_t)(n0 + c) * ld],
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (done + 1 < i) fetch_l(done + 1);
      __syncthreads();
      mma64_kmajor<T>(As, PP, Bs, PB, acc);
    }
  }
  // X_i = DINV_i (S_i - acc)
  __syncthreads();
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = acc_row<T>(bm, q), col = acc_col(bn);
        Bs[row * PB + col] = sv[bm][bn][q] - acc[bm][bn][q];
      }
#pragma unroll
  for (int q = 0; q < 16; ++q) As[(cb + 4 * q) * PP + r] = dv[q];
  __syncthreads();
  zero_acc<T>(acc);
  mma64_kmajor<T>(As, PP, Bs, PB, acc);
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        __hip_atomic_store(&S[(size_t)(ri + acc_row<T>(bm, q)) + (size_t)(n0 + acc_col(bn)) * ld],
                           (T)acc[bm][bn][q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  publish_wait();  // every wave's X_i stores complete ...
  __syncthreads();
  if (threadIdx.x == 0)  // ... before (i + 1, ct) may read them
    __hip_atomic_store(prog, ((uint64_t)epoch << 32) | (uint64_t)(i + 1), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Transposed wavefront TRSM: X <- L^-T X in place on the predict buffer S,
// for the column tiles [ct0, ct0 + nct) (nct <= WAVE_NCT), all members in
// one launch.  Step t of a member is its block row i = nb - 1 - t (the sweep
// runs from the last block row up): workgroup (t, member, ct) accumulates
// L_ji^T X_j[:, ct] for every j > i as soon as step nb - 1 - j publishes,
// then forms X_i = DINV_i^T (X_i - acc) and publishes step t.  Operands as in
// k_trsm_rows with both L_ji and DINV_i transposed on their way into LDS
// (rows of the stored tile are the k index, read coalesced along k).
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_trsm_rows_t(Prob* const* probs, Sched rows, int nct,
                                                          int ct0, uint32_t* ticket,
                                                          uint32_t epoch) {
  constexpr int PA = TILE + 1;  // A operand, written row by row
  constexpr int PB = TILE + 1;  // B operand, written transposed
  __shared__ T As[TILE * PA];
  __shared__ T Bs[TILE * PB];
  __shared__ int s_val;
  typedef typename Acc<T>::type acc_t;
  const int tk = take_ticket(ticket, &s_val);
  if ((unsigned)tk >= gridDim.x) {  // counter not zero at launch: no block to take
    flag_members_stalled(probs, (rows.pre[1] - rows.pre[0]) / nct);  // step 0: every member
    return;
  }
  int local;
  const int t = sched_member(rows, tk, local);
  const int m = local / nct, lct = local - m * nct;
  const Prob& P = *probs[m];
  const int nb = P.Np / TILE, ct = ct0 + lct;
  if (ct >= P.Mp / TILE || t >= nb) return;
  const int i = nb - 1 - t;
  uint64_t* prog = P.progress_ct + lct;
  const auto L = gbl((const T*)P.L);
  T* S = (T*)P.S;
  const int64_t ld = P.Np;
  const int64_t ri = (int64_t)i * TILE, n0 = (int64_t)ct * TILE;
  const int r = threadIdx.x & 63, cb = threadIdx.x >> 6;
  T dv[16], la[16], sv[2][2][4];
  {
    const auto D = gbl(P.dinv + (size_t)i * TILE * TILE);
    const auto Sg = gbl((const T*)P.S);
#pragma unroll
    for (int q = 0; q < 16; ++q) dv[q] = (T)D[r + (size_t)(cb + 4 * q) * TILE];
#pragma unroll
    for (int bm = 0; bm < 2; ++bm)
#pragma unroll
      for (int bn = 0; bn < 2; ++bn)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          sv[bm][bn][q] =
              Sg[(size_t)(ri + acc_row<T>(bm, q)) + (size_t)(n0 + acc_col(bn)) * ld];
  }
  // la[q] = L_ji(r, cb + 4q) = L(64 j + r, 64 i + cb + 4q): A(m, k) = L_ji(k, m)
  auto fetch_l = [&](int j) {
#pragma unroll
    for (int q = 0; q < 16; ++q)
      la[q] = L[(size_t)((int64_t)j * TILE + r) + (size_t)(ri + cb + 4 * q) * ld];
  };
  if (t > 0) fetch_l(nb - 1);
  acc_t acc[2][2];
  zero_acc<T>(acc);
  int done = 0, ready = 0;
  while (done < t) {
    if (done >= ready) {
      __syncthreads();  // everyone has read the previous s_val
      if (threadIdx.x == 0) s_val = wait_progress(prog, epoch, done + 1, P.info);
      __syncthreads();
      ready = min(__builtin_amdgcn_readfirstlane(s_val), t);
    }
    for (; done < ready; ++done) {
      const int64_t rj = (int64_t)(nb - 1 - done) * TILE;
      __syncthreads();  // the previous product is done with LDS
#pragma unroll
      for (int q = 0; q < 16; ++q) As[r * PA + cb + 4 * q] = la[q];
      // B[kk][n] = X_j(kk, n), published in this launch: coherent loads
#pragma unroll 4
      for (int q = 0; q < 16; ++q) {
        const int c = cb + 4 * q;
        Bs[r * PB + c] = __hip_atomic_load(&S[(size_t)(rj + r) + (size_t)(n0 + c) * ld],
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (done + 1 < t) fetch_l(nb - 2 - done);
      __syncthreads();
      mma64_kmajor<T>(As, PA, Bs, PB, acc);
    }
  }
  // X_i = DINV_i^T (X_i - acc)
  __syncthreads();
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = acc_row<T>(bm, q), col = acc_col(bn);
        Bs[row * PB + col] = sv[bm][bn][q] - acc[bm][bn][q];
      }
#pragma unroll
  for (int q = 0; q < 16; ++q) As[r * PA + cb + 4 * q] = dv[q];  // A(m, k) = DINV_i(k, m)
  __syncthreads();
  zero_acc<T>(acc);
  mma64_kmajor<T>(As, PA, Bs, PB, acc);
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        __hip_atomic_store(&S[(size_t)(ri + acc_row<T>(bm, q)) + (size_t)(n0 + acc_col(bn)) * ld],
                           (T)acc[bm][bn][q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  publish_wait();  // every wave's X_i stores complete ...
  __syncthreads();
  if (threadIdx.x == 0)  // ... before step t + 1 may read them
    __hip_atomic_store(prog, ((uint64_t)epoch << 32) | (uint64_t)(t + 1), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// S <- L^-T S for every member's Mp right-hand sides (Mp <= maxMp): one
// wavefront launch per group of WAVE_NCT column tiles.
void launch_trsm_t(const Batch& bt, int prec, int maxMp, hipStream_t s) {
  const int nb = bt.maxNp / TILE;
  const int mt = maxMp / TILE;
  for (int ct0 = 0; ct0 < mt; ct0 += WAVE_NCT) {
    const int nct = std::min(WAVE_NCT, mt - ct0);
    Sched rows;
    rows.count = nb;
    int tot = 0;
    for (int t = 0; t < nb; ++t) {
      rows.pre[t] = tot;
      for (int m = 0; m < bt.count; ++m)
        if (bt.np(m) / TILE > t) tot += nct;
    }
    rows.pre[nb] = tot;
    if (tot == 0) return;
    const bool on = timer_on(TC_TRSM);
    double fl = 0.0;
    if (on) {
      for (int m = 0; m < bt.count; ++m) {
        const double np = bt.np(m);
        fl += np * np * (double)nct * TILE;
      }
      timer_begin(TC_TRSM, s);
    }
    const uint32_t ep = next_epoch();
    if (prec == PGP_PREC_FP64)
      hipLaunchKernelGGL(k_trsm_rows_t<double>, dim3(tot), dim3(NTHREADS), 0, s, bt.dprobs, rows,
                         nct, ct0, bt.ticket, ep);
    else
      hipLaunchKernelGGL(k_trsm_rows_t<float>, dim3(tot), dim3(NTHREADS), 0, s, bt.dprobs, rows,
                         nct, ct0, bt.ticket, ep);
    if (on) timer_end(TC_TRSM, s, fl, 0.0);
  }
}

static bool wavefront_trsm() {
  static const bool v = [] {
    const char* e = getenv("PGP_TRSM_WAVEFRONT");
    return !(e && e[0] == '0');
  }();
  return v;
}

// Super-steps of G block rows.
// One launch per super-step q of block rows SS_q = [Gq, Gq + G), with
// P = SS_{q-1} the super-step finished by the previous launch:
//   solve workgroup (column tile ct), for k in SS_q in order:
//     X_k = DINV_k (S_k - sum_{c in [G(q-1), k)} L_kc X_c)   (left-looking)
//   update workgroups (row block i >= G(q+1), column tile ct):
//     S_i -= sum_{c in P} L_ic X_c                          (rank 64 G)
// so every later row block receives each finished super-step in one
// rank-64G product: S is read and written once per super-step, not once per
// block row.  A workgroup writes only its own rows and reads rows finished
// in earlier launches (or by itself), so workgroups of one launch never
// race.  lower_rhs: S is lower triangular (identity-like right-hand side):
// the blocks X_c[:, ct] with c < ct are zero and are skipped.
constexpr int TRSM_G = 4;

// acc += sum_{c in [c0, c1)} L(r0.., 64c..) S(64c.., n0..): MFMA tile
// products with the next chunk fetched into registers during the current.
// Rows of S from block c_fresh on were written by this workgroup in this
// launch: they are read with coherent loads (past the CU's L1, which may
// still hold the values they replaced).
template <typename T>
__device__ __forceinline__ void trsm_acc(const T* L, int64_t ld, const T* S, int64_t lds, int64_t r0, int64_t n0,
                         int c0, int c1, int c_fresh, T* As, T* Bs,
                         typename Acc<T>::type (&acc)[2][2]) {
  constexpr int PP = KP<T>::value;  // A operand: conflict-free MFMA reads
  constexpr int PB = TILE + 1;       // B operand: written transposed, conflict-free stores
  if (c0 >= c1) return;
  const int m = threadIdx.x & 63, kb = threadIdx.x >> 6;
  T ra[16], rb[16];
  auto fetch = [&](int c) {
    const int64_t cc = (int64_t)c * TILE;
    const bool fresh = c >= c_fresh;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = kb + 4 * q;
      ra[q] = L[(size_t)(r0 + m) + (size_t)(cc + k) * ld];  // A(m, k)
      const T* sp = &S[(size_t)(cc + m) + (size_t)(n0 + k) * lds];  // B(m, k): row m, column k
      rb[q] = fresh ? __hip_atomic_load(sp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : *sp;
    }
  };
  fetch(c0);
  for (int c = c0; c < c1; ++c) {
    if (c > c0) __syncthreads();  // the previous chunk's products are done with LDS
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = kb + 4 * q;
      As[k * PP + m] = ra[q];
      Bs[m * PB + k] = rb[q];
    }
    __syncthreads();
    if (c + 1 < c1) fetch(c + 1);
    mma64_kmajor<T>(As, PP, Bs, PB, acc);
  }
  __syncthreads();  // LDS free for the caller
}

// Workgroup `u` of launch q for one right-hand side of nct column tiles:
// u < nct solves SS_q for column tile u; the others update one row block.
template <typename T>
__device__ __forceinline__ void trsm_ss(const T* L, int64_t ld, const double* dinv, int nb, int q, T* S,
                        int64_t lds, int nct, int u, int lower_rhs) {
  constexpr int PP = KP<T>::value;
  constexpr int PB = TILE + 1;
  __shared__ T As[TILE * PP];
  __shared__ T Bs[TILE * PB];
  typedef typename Acc<T>::type acc_t;
  constexpr int G = TRSM_G;
  const int p0 = q > 0 ? G * (q - 1) : 0;  // first column of P (or of SS_0)
  acc_t acc[2][2];
  if (u < nct) {
    const int ct = u;
    const int64_t n0 = (int64_t)ct * TILE;
    const int kend = min(G * q + G, nb);
    for (int k = G * q; k < kend; ++k) {
      if (lower_rhs && ct > k) continue;  // X_k[:, ct] stays zero
      const int64_t r0 = (int64_t)k * TILE;
      zero_acc<T>(acc);
      trsm_acc<T>(L, ld, S, lds, r0, n0, lower_rhs ? max(p0, ct) : p0, k, G * q, As, Bs, acc);
      // B = S_k - acc (rounded to T), A = DINV_k
#pragma unroll
      for (int bm = 0; bm < 2; ++bm)
#pragma unroll
        for (int bn = 0; bn < 2; ++bn)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = acc_row<T>(bm, r), col = acc_col(bn);
            Bs[row * PB + col] =
                S[(size_t)(r0 + row) + (size_t)(n0 + col) * lds] - acc[bm][bn][r];
          }
      load_block_kmajor<double, T>(As, PP, dinv + (size_t)k * TILE * TILE, TILE, 0, 0);
      __syncthreads();
      zero_acc<T>(acc);
      mma64_kmajor<T>(As, PP, Bs, PB, acc);
#pragma unroll
      for (int bm = 0; bm < 2; ++bm)
#pragma unroll
        for (int bn = 0; bn < 2; ++bn)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            S[(size_t)(r0 + acc_row<T>(bm, r)) + (size_t)(n0 + acc_col(bn)) * lds] =
                acc[bm][bn][r];
      publish_wait();   // X_k reaches L2 ...
      __syncthreads();  // ... before the next row of this super-step reads it back
    }
    return;
  }
  // update of row block i >= G(q+1) by P = [G(q-1), Gq)
  const int v = u - nct;
  const int ct = v % nct;
  const int i = G * (q + 1) + v / nct;
  if (q == 0 || i >= nb) return;
  const int c0 = lower_rhs ? max(p0, ct) : p0;
  if (c0 >= G * q) return;
  const int64_t r0 = (int64_t)i * TILE, n0 = (int64_t)ct * TILE;
  zero_acc<T>(acc);
  trsm_acc<T>(L, ld, S, lds, r0, n0, c0, G * q, G * q, As, Bs, acc);
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        T* p = &S[(size_t)(r0 + acc_row<T>(bm, r)) + (size_t)(n0 + acc_col(bn)) * lds];
        *p = *p - acc[bm][bn][r];
      }
}

// workgroups of launch q: nct solvers plus nct per row block below SS_q
static inline int trsm_units(int nb, int q, int nct) {
  const int rows = q > 0 ? std::max(0, nb - TRSM_G * (q + 1)) : 0;
  return nb > TRSM_G * q ? nct * (1 + rows) : 0;
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_trsm(Prob* const* probs, Sched sc, int q,
                                                   int lower_rhs) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  trsm_ss<T>((const T*)P.L, P.Np, P.dinv, P.Np / TILE, q, (T*)P.S, P.Np, P.Mp / TILE, bx,
             lower_rhs);
}

static void launch_trsm_lagged(const Batch& bt, int prec, int maxMp, int lower_rhs, hipStream_t s) {
  const int nb = bt.maxNp / TILE;
  const bool on = timer_on(TC_TRSM);
  for (int k = 0; k < nb; ++k) {
    // lower_rhs: column tiles ct > k of an identity-like right-hand side
    // stay zero, so only ct <= k are scheduled
    auto nct_of = [&](int m) {
      const int mp = bt.hm ? (bt.hm[m] + TILE - 1) / TILE * TILE : maxMp;
      return lower_rhs ? std::min(k + 1, mp / TILE) : mp / TILE;
    };
    sched_launch(
        bt,
        [&](int m) {
          const int nbm = bt.np(m) / TILE;
          return k < nbm ? (k == 0 ? 1 : nbm - k) * nct_of(m) : 0;
        },
        [&](Prob* const* dp, const Sched& sc, int n) {
          if (on) timer_begin(TC_TRSM, s);
          if (prec == PGP_PREC_FP64)
            hipLaunchKernelGGL(k_trsm_lagged<double>, dim3(n), dim3(NTHREADS), 0, s, dp, sc, k, lower_rhs);
          else
            hipLaunchKernelGGL(k_trsm_lagged<float>, dim3(n), dim3(NTHREADS), 0, s, dp, sc, k, lower_rhs);
          if (on) {
            // step k of an N^2 M-flop solve: the update of rows k.. by block
            // k - 1 (2 * 64 per entry) and the 64-row triangle (64^2 per column)
            double fl = 0.0;
            const int off = (int)(dp - bt.dprobs);
            for (int m = off; m < off + sc.count; ++m) {
              const int nbm = bt.np(m) / TILE;
              if (k >= nbm) continue;
              const double cols = (double)nct_of(m) * TILE;
              fl += cols * (TILE * TILE + (k > 0 ? 2.0 * TILE * TILE * (nbm - k) : 0.0));
            }
            timer_end(TC_TRSM, s, fl, 0.0);
          }
        });
  }
}

// Right-hand sides of at least this many column tiles take the super-step
// schedule.
constexpr int TRSM_WIDE = 16;

void launch_trsm(const Batch& bt, int prec, int maxMp, int lower_rhs, hipStream_t s) {
  const int nct = maxMp / TILE;
  if (!lower_rhs && bt.ticket && nct <= WAVE_NCT && bt.maxNp / TILE <= MAXB && wavefront_trsm()) {
    // row-major schedule: row i holds the members with more than i block
    // rows (a prefix: members are sorted by size), nct workgroups each
    const int nb = bt.maxNp / TILE;
    Sched rows;
    rows.count = nb;
    int tot = 0;
    double fl = 0.0;
    for (int i = 0; i < nb; ++i) {
      rows.pre[i] = tot;
      for (int m = 0; m < bt.count; ++m)
        if (bt.np(m) / TILE > i) tot += nct;
    }
    rows.pre[nb] = tot;
    if (tot == 0) return;
    const bool on = timer_on(TC_TRSM);
    if (on) {
      for (int m = 0; m < bt.count; ++m) {
        const double np = bt.np(m);
        const double mp = bt.hm ? (bt.hm[m] + TILE - 1) / TILE * TILE : maxMp;
        fl += np * np * mp;
      }
      timer_begin(TC_TRSM, s);
    }
    const uint32_t ep = next_epoch();
    if (prec == PGP_PREC_FP64)
      hipLaunchKernelGGL(k_trsm_rows<double>, dim3(tot), dim3(NTHREADS), 0, s, bt.dprobs, rows,
                         nct, bt.ticket, ep);
    else
      hipLaunchKernelGGL(k_trsm_rows<float>, dim3(tot), dim3(NTHREADS), 0, s, bt.dprobs, rows,
                         nct, bt.ticket, ep);
    if (on) timer_end(TC_TRSM, s, fl, 0.0);
    return;
  }
  if (maxMp / TILE < TRSM_WIDE) {
    launch_trsm_lagged(bt, prec, maxMp, lower_rhs, s);
    return;
  }
  const int nb = bt.maxNp / TILE;
  const int nq = (nb + TRSM_G - 1) / TRSM_G;
  const bool on = timer_on(TC_TRSM);
  auto nct_of = [&](int m) {
    const int mp = bt.hm ? (bt.hm[m] + TILE - 1) / TILE * TILE : maxMp;
    return mp / TILE;
  };
  for (int q = 0; q < nq; ++q) {
    sched_launch(
        bt, [&](int m) { return trsm_units(bt.np(m) / TILE, q, nct_of(m)); },
        [&](Prob* const* dp, const Sched& sc, int n) {
          if (on) timer_begin(TC_TRSM, s);
          if (prec == PGP_PREC_FP64)
            hipLaunchKernelGGL(k_trsm<double>, dim3(n), dim3(NTHREADS), 0, s, dp, sc, q, lower_rhs);
          else
            hipLaunchKernelGGL(k_trsm<float>, dim3(n), dim3(NTHREADS), 0, s, dp, sc, q, lower_rhs);
          if (on) {
            // super-step q of an N^2 M-flop solve: its rows' left-looking
            // products and triangles, and the rank-64G update of the rows
            // below by the previous super-step (full right-hand side)
            constexpr double T3 = 2.0 * TILE * TILE * TILE;
            double fl = 0.0;
            const int off = (int)(dp - bt.dprobs);
            for (int m = off; m < off + sc.count; ++m) {
              const int nbm = bt.np(m) / TILE, nct = nct_of(m);
              if (nbm <= TRSM_G * q) continue;
              const int p0 = q > 0 ? TRSM_G * (q - 1) : 0;
              for (int k = TRSM_G * q; k < std::min(nbm, TRSM_G * q + TRSM_G); ++k)
                fl += nct * (T3 * (k - p0) + T3 / 2);
              if (q > 0)
                fl += (double)nct * std::max(0, nbm - TRSM_G * (q + 1)) * T3 * TRSM_G;
            }
            timer_end(TC_TRSM, s, fl, 0.0);
          }
        });
  }
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_trsm_raw(const T* L, int64_t ld, const double* dinv,
                                                       int nb, int q, T* S, int64_t lds, int nct,
                                                       int lower_rhs) {
  trsm_ss<T>(L, ld, dinv, nb, q, S, lds, nct, blockIdx.x, lower_rhs);
}

int launch_trsm_raw_dinv(int prec, const void* L, int64_t N, int64_t ld, const double* dinv,
                         void* S, int64_t M, int64_t lds, int lower_rhs, hipStream_t s) {
  const int nb = (int)(N / TILE), nct = (int)(M / TILE);
  if (nct < TRSM_WIDE) {
    for (int k = 0; k < nb; ++k) {
      dim3 g(k == 0 ? 1 : nb - k, (unsigned)nct);
      if (prec == PGP_PREC_FP64)
        hipLaunchKernelGGL(k_trsm_lagged_raw<double>, g, dim3(NTHREADS), 0, s, (const double*)L,
                           ld, dinv, k, (double*)S, lds, lower_rhs);
      else
        hipLaunchKernelGGL(k_trsm_lagged_raw<float>, g, dim3(NTHREADS), 0, s, (const float*)L, ld,
                           dinv, k, (float*)S, lds, lower_rhs);
    }
    return 0;
  }
  const int nq = (nb + TRSM_G - 1) / TRSM_G;
  for (int q = 0; q < nq; ++q) {
    const dim3 g((unsigned)trsm_units(nb, q, nct));
    if (prec == PGP_PREC_FP64)
      hipLaunchKernelGGL(k_trsm_raw<double>, g, dim3(NTHREADS), 0, s, (const double*)L, ld, dinv,
                         nb, q, (double*)S, lds, nct, lower_rhs);
    else
      hipLaunchKernelGGL(k_trsm_raw<float>, g, dim3(NTHREADS), 0, s, (const float*)L, ld, dinv,
                         nb, q, (float*)S, lds, nct, lower_rhs);
  }
  return 0;
}

}  // namespace pgp
