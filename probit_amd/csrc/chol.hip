// Blocked right-looking Cholesky A = L L^T (lower), replacing B.cholesky
// (probit/implicit/Laplace.py:24, probit/implicit/VB.py:10,25) and the LU
// solve of the Newton system (probit/implicit/solvers.py:24), whose matrix
// -(I + W K) is factored here in the symmetric form B = I + W^1/2 K W^1/2.
//
// Schedule for every batch member at once (c0 = 64k), with the block
// columns grouped into super-steps of G columns:
//   U_k  : left-looking update A_ik -= sum_c L_ic L_kc^T of column k (i >= k)
//          by the earlier columns c of k's super-step, 64x64 MFMA tiles;
//   F_k  : factor and invert diagonal block k (one workgroup per member);
//   P_k  : panel  L_ik = A_ik L_kk^-T for all row blocks i > k, as an MFMA
//          product with the stored inverse of L_kk;
//   O    : after the super-step's last column, the outer update of every
//          lower tile right of it with rank 64 G (FP64 / FP32 operands, or
//          BF16 operands with FP32 accumulation).
// A diagonal block is factored by one workgroup in FP64 together with its
// inverse (blocked by 16: each 16-column panel by one wave, one row per
// lane; the trailing updates and the inverse's off-diagonal blocks on FP64
// MFMA); the FP64 inverses (DINV) serve the panel products and the
// triangular solves.
#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

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

struct FactorOut {
  void* L;          // factor storage (T), ld
  int64_t ld;
  double* dinv;     // 64x64 FP64 inverse of this block, column-major
  double* logdiag;  // 64 entries for this block (may be null)
  double* diagval;  // 64 entries (may be null)
  int32_t* info;
};

// 1 / sqrt(p) in FP64: hardware estimate refined by two Newton steps
// r <- r (3 - p r^2) / 2 (each step doubles the correct bits).
__device__ __forceinline__ double rsqrt_f64(double p) {
  double r = __builtin_amdgcn_rsq(p);
  r = r * (1.5 - 0.5 * p * r * r);
  return r * (1.5 - 0.5 * p * r * r);
}

// Broadcast lane `lane`'s double to the whole wave (through scalar registers).
__device__ __forceinline__ double bcast(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// 16x16x16 FP64 MFMA product on LDS operands ([row][65] arrays), one wave:
// acc(i, j) += sum_k A(ar + i, ac + k) * B(k, j), with B(k, j) = Bm(br + k, bc + j)
// (bt = 0) or Bm(br + j, bc + k) (bt = 1).  acc[r] holds D(row (l >> 4) + 4r,
// col l & 15) (lane l).
__device__ __forceinline__ void mma16(const double (*A)[TILE + 1], int ar, int ac,
                                      const double (*Bm)[TILE + 1], int br, int bc, int bt,
                                      doublex4& acc) {
  const int l = threadIdx.x & 63;
  const int lr = l & 15, lk = l >> 4;
#pragma unroll
  for (int kk = 0; kk < 16; kk += 4) {
    const double a = A[ar + lr][ac + kk + lk];
    const double b = bt ? Bm[br + lr][bc + kk + lk] : Bm[br + kk + lk][bc + lr];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  }
}
// mma16 on row-major LDS arrays of row pitch PA (A) and PB (Bm).
template <int PA, int PB>
__device__ __forceinline__ void mma16p(const double* A, int ar, int ac, const double* Bm, int br,
                                       int bc, int bt, doublex4& acc) {
  const int l = threadIdx.x & 63;
  const int lr = l & 15, lk = l >> 4;
#pragma unroll
  for (int kk = 0; kk < 16; kk += 4) {
    const double a = A[(ar + lr) * PA + ac + kk + lk];
    const double b = bt ? Bm[(br + lr) * PB + bc + kk + lk] : Bm[(br + kk + lk) * PB + bc + lr];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  }
}

// LDS of the diagonal-block factor: the block A ([64][65] doubles) and the
// four 16x16 diagonal blocks of its inverse ([4][16][17]); the inverse's
// off-diagonal blocks live, transposed, in A's upper triangle.
constexpr int XD_PITCH = 17;
constexpr int XD_BLK = 16 * XD_PITCH;
constexpr int FACTOR_LDS = (TILE * (TILE + 1) + 4 * XD_BLK) * 8;

// Factor the 16-column panel A(o.., o..o+15) of a 64x64 block in place: one
// wave, lane l holds row o + l (16 values in registers; lanes past the
// block's last row idle).  Values another row needs travel through scalar
// registers (readlane), never through LDS; the pivot chain stays short
// because row j+1's diagonal entry only needs its own row.  The rows below
// the diagonal sub-block are finished here too, so no inverse is needed on
// the way.  dval / rinv receive L(o+j, o+j) and its reciprocal square root.
__device__ __forceinline__ void panel16(double (*A)[TILE + 1], int o, double* dval,
                                        double* rinv, int32_t* info) {
  const int l = threadIdx.x & 63;
  const bool valid = l < TILE - o;
  double a[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) a[m] = valid ? A[o + l][o + m] : 0.0;
  double myri = 0.0;
  bool bad = false;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const double p = bcast(a[j], j);
    const double ri = rsqrt_f64(p);
    bad |= !(p > 0.0);
    myri = (l == j) ? ri : myri;
    const double lj = (l > j) ? a[j] * ri : (l == j ? p * ri : 0.0);
    a[j] = lj;
    // a(l, m) -= l(l, j) l(m, j) for the later columns m, the next pivot's
    // column first
#pragma unroll
    for (int m = j + 1; m < 16; ++m) a[m] -= lj * bcast(lj, m);
  }
  if (l == 0 && bad && info) atomicOr(info, PGP_INFO_NOT_PD);
  if (valid) {
#pragma unroll
    for (int m = 0; m < 16; ++m)
      if (m <= l) A[o + l][o + m] = a[m];
  }
  if (l < 16) {
    dval[o + l] = a[l];
    rinv[o + l] = myri;
  }
}

// X(o.., o..) = inverse of the factored 16x16 diagonal sub-block at (o, o)
// (zero above the diagonal).  One wave; lane c runs forward substitution on
// column c of I, reading L from LDS (one address per step: a broadcast).
__device__ __forceinline__ void invert16(const double (*A)[TILE + 1], double* Xb, int o,
                                         const double* rinv) {
  const int c = threadIdx.x & 15;
  double x[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    double acc0 = (r == c) ? 1.0 : 0.0, acc1 = 0.0;
#pragma unroll
    for (int k = 0; k + 1 < r; k += 2) {
      acc0 -= A[o + r][o + k] * x[k];
      acc1 -= A[o + r][o + k + 1] * x[k + 1];
    }
    if (r & 1) acc0 -= A[o + r][o + r - 1] * x[r - 1];
    x[r] = (acc0 + acc1) * rinv[o + r];
  }
  if ((threadIdx.x & 63) < 16) {
#pragma unroll
    for (int r = 0; r < 16; ++r) Xb[r * XD_PITCH + c] = (r >= c) ? x[r] : 0.0;
  }
}

// Blocked factor + inverse of a 64x64 SPD block held in LDS (`A`, [row][65]).
// Per 16-wide panel s: one wave factors the panel (panel16), then FP64 MFMA
// products apply it to the trailing blocks, A_rc -= L_rs L_cs^T.  Then each
// wave inverts one diagonal sub-block, and the off-diagonal blocks of
// X = L^-1 follow by block distance: X_rs = -X_rr sum_{k=s}^{r-1} L_rk X_ks.
// X's diagonal blocks go to Xd ([4][16][17]); X_rs (r > s) is kept
// transposed in A's upper block (s, r), free once the factor is done, after
// serving as the scratch of T_rs = sum_k L_rk X_ks.
// L (lower) goes to global storage as T, X to o.dinv (FP64).
// `published(...)` runs (all threads) once the inverse's stores are issued,
// before L's: a caller that hands DINV to other workgroups publishes there.
struct NoPublish {
  __device__ void operator()() const {}
};
template <typename T, typename Pub = NoPublish>
__device__ __forceinline__ void factor_block_blocked(double (*A)[TILE + 1], double* Xd, int64_t r0,
                                     const FactorOut& o, Pub published = Pub()) {
  __shared__ double dval[TILE], rinv[TILE];
  const int t = threadIdx.x;
  const int wv = t >> 6, l = t & 63;
  const int orow = l >> 4, ocol = l & 15;  // D layout: row orow + 4r, col ocol
  __builtin_amdgcn_s_setprio(3);
  for (int s = 0; s < 4; ++s) {
    const int o16 = 16 * s;
    if (wv == 0) panel16(A, o16, dval, rinv, o.info);
    __syncthreads();
    if (s == 3) break;
    const int nrb = 3 - s;
    // trailing blocks (r, c), s < c <= r: A_rc -= L_rs L_cs^T
    for (int b = wv; b < nrb * (nrb + 1) / 2; b += 4) {
      int rr = 0;
      while ((rr + 1) * (rr + 2) / 2 <= b) ++rr;
      const int cc = b - rr * (rr + 1) / 2;
      const int r = s + 1 + rr, c = s + 1 + cc;
      doublex4 acc = {0.0, 0.0, 0.0, 0.0};
      mma16(A, 16 * r, o16, A, 16 * c, o16, 1, acc);
#pragma unroll
      for (int q = 0; q < 4; ++q) A[16 * r + orow + 4 * q][16 * c + ocol] -= acc[q];
    }
    __syncthreads();
  }
  invert16(A, Xd + wv * XD_BLK, 16 * wv, rinv);
  __syncthreads();
  // off-diagonal blocks of X by distance d; T goes to A's upper block (s, r)
  // and is replaced there by X_rs transposed: X_ks (k > s) is read back as
  // B(kk, j) = A(16 s + j, 16 k + kk)
  const double* Af = &A[0][0];
  for (int d = 1; d < 4; ++d) {
    const int nblk = 4 - d;
    if (wv < nblk) {
      const int r = d + wv, s = wv;
      doublex4 acc = {0.0, 0.0, 0.0, 0.0};
      mma16p<TILE + 1, XD_PITCH>(Af, 16 * r, 16 * s, Xd + s * XD_BLK, 0, 0, 0, acc);
      for (int k = s + 1; k < r; ++k)
        mma16p<TILE + 1, TILE + 1>(Af, 16 * r, 16 * k, Af, 16 * s, 16 * k, 1, acc);
#pragma unroll
      for (int q = 0; q < 4; ++q) A[16 * s + orow + 4 * q][16 * r + ocol] = acc[q];
    }
    __syncthreads();
    if (wv < nblk) {
      const int r = d + wv, s = wv;
      doublex4 acc = {0.0, 0.0, 0.0, 0.0};
      mma16p<XD_PITCH, TILE + 1>(Xd + r * XD_BLK, 0, 0, Af, 16 * s, 16 * r, 0, acc);
      // (the product has read all of T_rs before its result exists)
#pragma unroll
      for (int q = 0; q < 4; ++q) A[16 * s + ocol][16 * r + orow + 4 * q] = -acc[q];
    }
    __syncthreads();
  }
  __builtin_amdgcn_s_setprio(0);
  for (int e = t; e < TILE * TILE; e += NTHREADS) {
    const int c = e / TILE, r = e % TILE;  // column-major walk: coalesced stores
    double v = 0.0;
    if (c <= r)
      v = (r >> 4) == (c >> 4) ? Xd[(r >> 4) * XD_BLK + (r & 15) * XD_PITCH + (c & 15)] : A[c][r];
    // coherent: the column kernels hand DINV to other workgroups
    st_coherent(&o.dinv[r + c * TILE], v);
  }
  published();
  if (t < TILE) {
    if (o.logdiag) o.logdiag[t] = log(dval[t]);
    if (o.diagval) o.diagval[t] = dval[t];
  }
  T* L = (T*)o.L;
  for (int e = t; e < TILE * TILE; e += NTHREADS) {
    const int c = e / TILE, r = e % TILE;
    if (c <= r) L[(size_t)(r0 + r) + (size_t)(r0 + c) * o.ld] = (T)A[r][c];
  }
}

// ---- F: first diagonal block -------------------------------------------------
template <typename T>
__device__ void factor_first(const T* A, int64_t ld, const FactorOut& o) {
  __shared__ double blk[TILE][TILE + 1];
  __shared__ double xd[4 * XD_BLK];
  const int m = threadIdx.x & 63, kb = threadIdx.x >> 6;
  for (int q = 0; q < 16; ++q) {
    const int c = kb + 4 * q;
    blk[m][c] = (double)A[(size_t)m + (size_t)c * ld];
  }
  __syncthreads();
  factor_block_blocked<T>(blk, xd, 0, o);
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_factor_first(Prob* const* probs) {
  const Prob& P = *probs[blockIdx.y];
  FactorOut o{P.L, P.Np, P.dinv, P.logdiag, P.diagval, P.info};
  factor_first<T>((const T*)P.L, P.Np, o);
}

// ---- P_k: panel rows below block k --------------------------------------------
// L_ik = A_ik DINV_k^T in the storage type: FP64 MFMA for an FP64 factor,
// FP32 MFMA (the FP64 inverse rounded to FP32) for an FP32 one, whose
// trailing updates carry FP32 rounding anyway.
template <typename T>
__device__ void panel_rows(T* A, int64_t ld, const double* dinv_k, int64_t c0, int64_t r0) {
  constexpr int PP = KP<T>::value;
  __shared__ T As[TILE * PP];
  __shared__ T Bs[TILE * PP];
  // As[p][m] = A(r0 + m, c0 + p); Bs[p][n] = DINV_k(n, p)
  load_block_kmajor<T, T>(As, PP, A, ld, r0, c0);
  {
    const int n = threadIdx.x & 63, pb = threadIdx.x >> 6;
    for (int q = 0; q < 16; ++q) {
      const int p = pb + 4 * q;
      Bs[p * PP + n] = (T)dinv_k[n + p * TILE];
    }
  }
  __syncthreads();
  typename Acc<T>::type acc[2][2];
  zero_acc<T>(acc);
  mma64_kmajor<T>(As, PP, Bs, PP, acc);
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        A[(size_t)(r0 + acc_row<T>(bm, r)) + (size_t)(c0 + acc_col(bn)) * ld] = (T)acc[bm][bn][r];
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_panel(Prob* const* probs, Sched sc, int k) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int64_t c0 = (int64_t)k * TILE;
  const int64_t r0 = c0 + (int64_t)TILE * (bx + 1);
  if (r0 >= P.Np) return;
  panel_rows<T>((T*)P.L, P.Np, P.dinv + (size_t)k * TILE * TILE, c0, r0);
}

// ---- S_k: trailing update; tile (k+1, k+1) then factors its block ----------
template <typename T, int BF16, int FUSE>
__device__ void syrk_tile(T* A, int64_t ld, int64_t c0, int64_t ri, int64_t rj, bool factor,
                          const FactorOut& o) {
  // one LDS arena: MMA operands, then (fused mode, tile (k+1, k+1) only) the
  // two FP64 blocks of the diagonal factorisation
  constexpr int PP = KP<T>::value;
  constexpr int OPER = BF16 ? 2 * TILE * 72 * 2 : 2 * TILE * PP * (int)sizeof(T);
  constexpr int BLK = FUSE ? FACTOR_LDS : 0;
  __shared__ __attribute__((aligned(16))) char arena[OPER > BLK ? OPER : BLK];
  typedef typename Acc<T>::type acc_t;
  acc_t acc[2][2];
  zero_acc<T>(acc);
  // the C tile is loaded first, so its latency overlaps the product
  T cpre[2][2][4];
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        cpre[bm][bn][r] = A[(size_t)(ri + acc_row<T>(bm, r)) + (size_t)(rj + acc_col(bn)) * ld];
  if constexpr (BF16) {
    unsigned short* si = (unsigned short*)arena;
    unsigned short* sj = si + TILE * 72;
    load_block_bf16_mmajor<T>(si, 72, A, ld, ri, c0);
    load_block_bf16_mmajor<T>(sj, 72, A, ld, rj, c0);
    __syncthreads();
    mma64_bf16(si, 72, sj, 72, acc);
  } else {
    T* si = (T*)arena;
    T* sj = si + TILE * PP;
    load_block_kmajor<T, T>(si, PP, A, ld, ri, c0);
    load_block_kmajor<T, T>(sj, PP, A, ld, rj, c0);
    __syncthreads();
    mma64_kmajor<T>(si, PP, sj, PP, acc);
  }
  if constexpr (!FUSE) factor = false;
  double (*blk)[TILE + 1] = (double (*)[TILE + 1])arena;
  if (factor) __syncthreads();  // operands are dead; the arena becomes the block
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn) {
      const int col = acc_col(bn);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = acc_row<T>(bm, r);
        const T v = cpre[bm][bn][r] - acc[bm][bn][r];
        if (factor)
          blk[row][col] = (double)v;
        else
          A[(size_t)(ri + row) + (size_t)(rj + col) * ld] = v;
      }
    }
  if constexpr (FUSE) {
    if (factor) {
      __syncthreads();
      factor_block_blocked<T>(blk, (double*)(blk + TILE), ri, o);
    }
  }
}

// The trailing updates of FP32 factors (outer update, column update, column
// kernel) stage their operands m-major for mma64_f32_mmajor; FP64 factors
// stay k-major and BF16 operands are m-major bf16.  All three kernels use
// the same form, so a tile's update is bitwise the same in each.
template <typename T, int BF16>
struct MmajorF32 {
  static constexpr bool value = std::is_same<T, float>::value && !BF16;
};

// k index of the q-th of the 16 operand values a thread (row m, group kb)
// fetches per 64-column chunk: 16 consecutive k for the m-major FP32 form,
// every fourth k otherwise.  Either way one fetch instruction reads one
// column's 64 consecutive rows.
template <bool MM>
__device__ __forceinline__ int operand_k(int kb, int q) {
  return MM ? 16 * kb + q : kb + 4 * q;
}

// d[m * pitch + 16 kb + q] = v[q], q = 0..15, as four 16-byte stores
// (pitch a multiple of 4).
template <typename T>
__device__ __forceinline__ void stash_f32_mmajor(T* d, int pitch, const T (&v)[16], int m, int kb) {
  static_assert(std::is_same<T, float>::value, "m-major FP32 operands only");
#pragma unroll
  for (int j = 0; j < 4; ++j)
    *(floatx4*)&d[m * pitch + 16 * kb + 4 * j] =
        floatx4{v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]};
}

// d[m * pitch + 16 kb + q] = bf16(v[q]), q = 0..15, as two 16-byte stores
// (pitch a multiple of 8): the same LDS image as one 2-byte store per value.
template <typename T>
__device__ __forceinline__ void stash_bf16_mmajor(unsigned short* d, int pitch, const T (&v)[16],
                                                  int m, int kb) {
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    u16x8 w;
#pragma unroll
    for (int e = 0; e < 8; ++e) w[e] = f2bf((float)v[8 * j + e]);
    *(u16x8*)&d[m * pitch + 16 * kb + 8 * j] = w;
  }
}

// Workgroups of one launch are dealt to the 8 XCDs round-robin (workgroup g
// runs on XCD g % 8).  Renumber them so that each XCD gets one contiguous
// run of tiles: neighbouring tiles share panel rows, which then stay in that
// XCD's L2.
__device__ __forceinline__ int xcd_contiguous(int g, int total) {
  const int x = g & 7, q = total >> 3, r = total & 7;
  return x * q + (x < r ? x : r) + (g >> 3);
}

// C(ri.., rj..) -= sum over block columns c in [kc0, kc1) of
// A(ri.., 64c..) A(rj.., 64c..)^T for one 64x64 tile: the rank-64(kc1-kc0)
// update of a super-step, with the next 64-wide chunk of both panels loaded
// into registers while the current one is multiplied from LDS.  C is read
// and written once however long the K range.
// FUSE: the workgroup given factor = true (the tile that becomes the next
// diagonal block) keeps its updated tile in LDS and factors it at once.
// Requires kc1 > kc0 (both callers update by at least one block column).
// SH (BF16 factors, the outer update): the operand chunks come as BF16 from
// the super-step's panel copy (P.panel16, row r at (r - p16_row0) * pitch),
// the same values the stash would convert, so no conversion and half the
// operand bytes.
template <typename T, int BF16, int FUSE, bool SH = false>
__device__ __forceinline__ void syrk_tile_multi(T* A_, int64_t ld, int kc0, int kc1, int64_t ri, int64_t rj,
                                bool factor, const FactorOut& o,
                                const unsigned short* p16 = nullptr, int64_t p16_row0 = 0,
                                int p16_pitch = 0) {
  static_assert(!SH || BF16, "the panel copy holds BF16 operands");
  auto A = gbl(A_);
  constexpr bool MM = MmajorF32<T, BF16>::value;
  constexpr int PITCH = BF16 ? 72 : MM ? 68 : KP<T>::value;
  typedef typename std::conditional<BF16 != 0, unsigned short, T>::type LT;
  constexpr int OPER = 2 * TILE * PITCH * (int)sizeof(LT);
  constexpr int BLK = FUSE ? FACTOR_LDS : 0;
  __shared__ __attribute__((aligned(16))) char arena[OPER > BLK ? OPER : BLK];
  LT* sa = (LT*)arena;
  LT* sb = sa + TILE * PITCH;
  typedef typename Acc<T>::type acc_t;
  const int m = threadIdx.x & 63, kb = threadIdx.x >> 6;
  T ra[16], rb[16];
  acc_t acc[2][2];
  zero_acc<T>(acc);
  // the m-major FP32 and BF16 stagings take 16 consecutive k per thread
  constexpr bool CK = MM || BF16;
  u16x8 ua[2], ub[2];  // SH: a thread's 16 BF16 operand values of each row
  auto fetch = [&](int c) {
    if constexpr (SH) {
      const size_t kofs = (size_t)(c - kc0) * TILE + 16 * kb;
      const unsigned short* pa = p16 + (size_t)(ri + m - p16_row0) * p16_pitch + kofs;
      const unsigned short* pb = p16 + (size_t)(rj + m - p16_row0) * p16_pitch + kofs;
      ua[0] = *(const u16x8*)pa;
      ua[1] = *(const u16x8*)(pa + 8);
      ub[0] = *(const u16x8*)pb;
      ub[1] = *(const u16x8*)(pb + 8);
    } else {
      const int64_t c0 = (int64_t)c * TILE;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = operand_k<CK>(kb, q);
        ra[q] = A[(size_t)(ri + m) + (size_t)(c0 + k) * ld];
        rb[q] = A[(size_t)(rj + m) + (size_t)(c0 + k) * ld];
      }
    }
  };
  fetch(kc0);
  auto stash = [&](LT* da, LT* db) {
    if constexpr (SH) {
      *(u16x8*)&da[m * PITCH + 16 * kb] = ua[0];
      *(u16x8*)&da[m * PITCH + 16 * kb + 8] = ua[1];
      *(u16x8*)&db[m * PITCH + 16 * kb] = ub[0];
      *(u16x8*)&db[m * PITCH + 16 * kb + 8] = ub[1];
    } else if constexpr (MM) {
      stash_f32_mmajor(da, PITCH, ra, m, kb);
      stash_f32_mmajor(db, PITCH, rb, m, kb);
    } else if constexpr (BF16) {
      stash_bf16_mmajor(da, PITCH, ra, m, kb);
      stash_bf16_mmajor(db, PITCH, rb, m, kb);
    } else {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = kb + 4 * q;
        if constexpr (BF16) {
          da[m * PITCH + k] = f2bf((float)ra[q]);
          db[m * PITCH + k] = f2bf((float)rb[q]);
        } else {
          da[k * PITCH + m] = ra[q];
          db[k * PITCH + m] = rb[q];
        }
      }
    }
  };
  auto mma = [&](const LT* da, const LT* db) {
    if constexpr (BF16)
      mma64_bf16(da, PITCH, db, PITCH, acc);
    else if constexpr (MM)
      mma64_f32_mmajor(da, PITCH, db, PITCH, acc);
    else
      mma64_kmajor<T>(da, PITCH, db, PITCH, acc);
  };
  for (int c = kc0; c < kc1 - 1; ++c) {
    if (c > kc0) __syncthreads();  // the previous chunk's products are done with LDS
    stash(sa, sb);
    __syncthreads();
    fetch(c + 1);
    mma(sa, sb);
  }
  if (kc1 - 1 > kc0) __syncthreads();
  stash(sa, sb);
  __syncthreads();
  // the C tile is loaded during the last chunk's product (the fetch
  // registers are free then), so its latency overlaps the product
  T cpre[2][2][4];
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        cpre[bm][bn][r] = A[(size_t)(ri + acc_row<T>(bm, r)) + (size_t)(rj + acc_col(bn)) * ld];
  mma(sa, sb);
  if constexpr (!FUSE) factor = false;
  double (*blk)[TILE + 1] = (double (*)[TILE + 1])arena;
  if (factor) __syncthreads();  // operands are dead; the arena becomes the block
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn) {
      const int col = acc_col(bn);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = acc_row<T>(bm, r);
        const T v = cpre[bm][bn][r] - (T)acc[bm][bn][r];
        if (factor)
          blk[row][col] = (double)v;
        else
          A[(size_t)(ri + row) + (size_t)(rj + col) * ld] = v;
      }
    }
  if constexpr (FUSE) {
    if (factor) {
      __syncthreads();
      factor_block_blocked<T>(blk, (double*)(blk + TILE), ri, o);
    }
  }
}

// Left-looking update of block column j inside its super-step [s0, ...):
// tiles (i, j), i >= j, receive A_ij -= sum_{c in [s0, j)} L_ic L_jc^T in one
// product (the same arithmetic as the column kernel's update), just before
// block j is factored and its panel formed.  FUSE: the workgroup of tile
// (j, j) factors that block at once.
template <typename T, int BF16, int FUSE>
__global__ __launch_bounds__(NTHREADS) void k_col_update(Prob* const* probs, Sched sc, int j,
                                                         int s0) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int nb = P.Np / TILE;
  const int i = j + bx;
  if (i >= nb) return;
  const int c1 = j * TILE;
  const FactorOut o{P.L, P.Np, P.dinv + (size_t)j * TILE * TILE,
                    P.logdiag ? P.logdiag + c1 : nullptr, P.diagval ? P.diagval + c1 : nullptr,
                    P.info};
  syrk_tile_multi<T, BF16, FUSE>((T*)P.L, P.Np, s0, j, (int64_t)i * TILE, (int64_t)j * TILE,
                                 FUSE && i == j, o);
}

// Tile t of a Tk x Tk lower triangle in blocked order: super-tiles of 8 x 8
// tiles, super-rows in order, the super-tiles of a super-row left to right
// (the diagonal one as a lower triangle), row-major inside.  Consecutive
// tiles (those one XCD runs together) then share a few 64-row panels, which
// stay in that XCD's L2.  Super-rows before I hold 32 I^2 + 4 I tiles.
__device__ __forceinline__ void blocked_tile(int t, int Tk, int& ti, int& tj) {
  int I = (int)((sqrt(16.0 + 128.0 * (double)t) - 4.0) / 64.0);
  while (32 * (I + 1) * (I + 1) + 4 * (I + 1) <= t) ++I;
  while (32 * I * I + 4 * I > t) --I;
  const int u = t - (32 * I * I + 4 * I);
  const int rI = min(8, Tk - 8 * I);
  const int off = rI * 8 * I;
  if (u < off) {
    const int J = u / (rI * 8), v = u - J * rI * 8;
    ti = 8 * I + (v >> 3);
    tj = 8 * J + (v & 7);
  } else {
    int a, c;
    lower_tile_c(u - off, a, c);
    ti = 8 * I + a;
    tj = 8 * I + c;
  }
}

// Outer update closing a super-step [s0, s_end): every lower tile (i, j),
// s_end <= j <= i, gets A_ij -= sum_{c in [s0, s_end)} L_ic L_jc^T.
template <typename T, int BF16, int FUSE>
__global__ __launch_bounds__(NTHREADS) void k_syrk_outer(Prob* const* probs, Sched sc, int s0,
                                                         int s_end) {
  int bx;
  const int g = xcd_contiguous(blockIdx.x, gridDim.x);
  const Prob& P = *probs[sched_member(sc, g, bx)];
  const int nb = P.Np / TILE;
  const int Tk = nb - s_end;
  if (Tk <= 0 || bx >= Tk * (Tk + 1) / 2) return;
  int ti, tj;
  blocked_tile(bx, Tk, ti, tj);
  const int c1 = s_end * TILE;
  const FactorOut o{P.L, P.Np, P.dinv + (size_t)s_end * TILE * TILE,
                    P.logdiag ? P.logdiag + c1 : nullptr, P.diagval ? P.diagval + c1 : nullptr,
                    P.info};
  // BF16 factors read the operands from the panel copy (k_panel_bf16)
  syrk_tile_multi<T, BF16, FUSE, BF16 != 0>(
      (T*)P.L, P.Np, s0, s_end, (int64_t)(s_end + ti) * TILE, (int64_t)(s_end + tj) * TILE,
      FUSE && ti == 0 && tj == 0, o, P.panel16, (int64_t)s_end * TILE, P.panel16_pitch);
}

// BF16 copy of the panel columns [s0, s_end) of the rows below the
// super-step, for its outer update: one workgroup per 64-row block, each
// thread converting 16 consecutive columns of one row per block column.
// The conversion is the one the update's stash applies (f2bf).
__global__ __launch_bounds__(NTHREADS) void k_panel_bf16(Prob* const* probs, Sched sc, int s0,
                                                         int s_end) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int nb = P.Np / TILE;
  if (bx >= nb - s_end) return;
  const int m = threadIdx.x & 63, kb = threadIdx.x >> 6;
  const int64_t r = (int64_t)(s_end + bx) * TILE + m;
  const auto A = gbl((const float*)P.L);
  unsigned short* dst = P.panel16 + (size_t)(r - (int64_t)s_end * TILE) * P.panel16_pitch;
  for (int c = s0; c < s_end; ++c) {
    float v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q)
      v[q] = A[(size_t)r + (size_t)((int64_t)c * TILE + 16 * kb + q) * P.Np];
    stash_bf16_mmajor(dst + (size_t)(c - s0) * TILE, 0, v, 0, kb);
  }
}

// ---- one launch per block column ---------------------------------------------
// C = A(ri.., rj..) - sum_{c in [kc0, kc1)} A(ri.., 64c..) A(rj.., 64c..)^T,
// computed as the trailing updates are (T / BF16 operands on MFMA, result
// rounded to T), then held in LDS as FP64: out[r][c] (kmajor = false) or
// out[c][r] (kmajor = true).  `arena` holds the operands first.
// COH: the operand columns were written in this launch by other workgroups
// (merged column kernel): they are read with coherent loads, and `hook(c)`
// (called by every thread before column c's operands are fetched) waits
// for them.  The arithmetic is the same either way.
struct NoHook {
  __device__ void operator()(int) const {}
};
// SAME: the diagonal tile (ri == rj), whose two operand rows are the same:
// fetched and staged once.
template <typename T, int BF16, typename OT, bool SAME = false, bool COH = false,
          typename Hook = NoHook>
__device__ __forceinline__ void update_tile_to_lds(const T* A_, int64_t ld, int kc0, int kc1, int64_t ri,
                                   int64_t rj, char* arena, OT* out, int opitch, bool kmajor,
                                   Hook hook = Hook()) {
  auto A = gbl(A_);
  constexpr bool MM = MmajorF32<T, BF16>::value;
  constexpr int PITCH = BF16 ? 72 : MM ? 68 : KP<T>::value;
  typedef typename std::conditional<BF16 != 0, unsigned short, T>::type LT;
  typedef typename Acc<T>::type acc_t;
  LT* sa = (LT*)arena;
  LT* sb = SAME ? sa : sa + TILE * PITCH;
  const int m = threadIdx.x & 63, kb = threadIdx.x >> 6;
  acc_t acc[2][2];
  zero_acc<T>(acc);
  if (kc1 > kc0) {
    T ra[16], rb[16];
    auto fetch = [&](int c) {
      hook(c);
      const int64_t c0 = (int64_t)c * TILE;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = operand_k<MM>(kb, q);
        if constexpr (COH) {
          ra[q] = ld_coh(&A[(size_t)(ri + m) + (size_t)(c0 + k) * ld]);
          if constexpr (!SAME) rb[q] = ld_coh(&A[(size_t)(rj + m) + (size_t)(c0 + k) * ld]);
        } else {
          ra[q] = A[(size_t)(ri + m) + (size_t)(c0 + k) * ld];
          if constexpr (!SAME) rb[q] = A[(size_t)(rj + m) + (size_t)(c0 + k) * ld];
        }
      }
    };
    fetch(kc0);
    for (int c = kc0; c < kc1; ++c) {
      if (c > kc0) __syncthreads();
      if constexpr (MM) {
        stash_f32_mmajor(sa, PITCH, ra, m, kb);
        if constexpr (!SAME) stash_f32_mmajor(sb, PITCH, rb, m, kb);
      } else {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int k = kb + 4 * q;
          if constexpr (BF16) {
            sa[m * PITCH + k] = f2bf((float)ra[q]);
            if constexpr (!SAME) sb[m * PITCH + k] = f2bf((float)rb[q]);
          } else {
            sa[k * PITCH + m] = ra[q];
            if constexpr (!SAME) sb[k * PITCH + m] = rb[q];
          }
        }
      }
      __syncthreads();
      if (c + 1 < kc1) fetch(c + 1);
      if constexpr (BF16)
        mma64_bf16(sa, PITCH, sb, PITCH, acc);
      else if constexpr (MM)
        mma64_f32_mmajor(sa, PITCH, sb, PITCH, acc);
      else
        mma64_kmajor<T>(sa, PITCH, sb, PITCH, acc);
    }
    __syncthreads();  // operands are dead: `out` may overlay them
  }
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn) {
      const int col = acc_col(bn);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = acc_row<T>(bm, r);
        const T v = A[(size_t)(ri + row) + (size_t)(rj + col) * ld] - (T)acc[bm][bn][r];
        if (kmajor)
          out[col * opitch + row] = (OT)v;
        else
          out[row * opitch + col] = (OT)v;
      }
    }
  __syncthreads();
}

// LDS of the column kernels, used in turn by a tile update's operands, its
// FP64 result with the diagonal factor's inverse blocks (FACTOR_LDS), or a
// row tile's panel-product operands.
template <typename T, int BF16>
constexpr int col_arena_bytes() {
  constexpr int PP = KP<T>::value;
  constexpr int PITCH = BF16 ? 72 : MmajorF32<T, BF16>::value ? 68 : PP;
  constexpr int oper = 2 * TILE * PITCH * (BF16 ? 2 : (int)sizeof(T));
  constexpr int prod = 2 * TILE * PP * (int)sizeof(T);
  constexpr int a = oper > prod ? oper : prod;
  return a > FACTOR_LDS ? a : FACTOR_LDS;
}

// Column j of every member in one launch.  Tickets [0, nd) are the diagonal
// workgroups of members 0..nd-1 (members are sorted by size, so these are
// the members with more than j block columns); the rest are row workgroups
// listed by `rows`.  Within the super-step [s0, ...), column j has not yet
// received the updates of columns s0..j-1; each workgroup applies them to
// its own tile (left-looking, rank 64 (j - s0)).  The diagonal workgroup
// then factors and inverts block j and publishes; a row workgroup waits for
// its member's publication and forms L_ij = C_ij DINV_j^T (FP64 MFMA).
template <typename T, int BF16>
__global__ __launch_bounds__(NTHREADS) void k_column(Prob* const* probs, Sched rows, int nd,
                                                     uint32_t* ticket, uint32_t epoch, int j,
                                                     int s0) {
  __shared__ __attribute__((aligned(16))) char arena_b[col_arena_bytes<T, BF16>()];
  double (*arena)[TILE + 1] = (double (*)[TILE + 1])arena_b;
  double* xd = (double*)arena_b + TILE * (TILE + 1);
  __shared__ int s_val;
  const int t = take_ticket(ticket, &s_val);
  if ((unsigned)t >= gridDim.x) {  // counter not zero at launch: nothing to take
    flag_members_stalled(probs, nd);  // the diagonal workgroups' members: every member
    return;
  }
  const int64_t c0 = (int64_t)j * TILE;
  if (t < nd) {
    const Prob& P = *probs[t];
    update_tile_to_lds<T, BF16, double, true>((const T*)P.L, P.Np, s0, j, c0, c0, (char*)arena,
                                        &arena[0][0], TILE + 1, false);
    FactorOut o{P.L, P.Np, P.dinv + (size_t)j * TILE * TILE,
                P.logdiag ? P.logdiag + c0 : nullptr, P.diagval ? P.diagval + c0 : nullptr,
                P.info};
    factor_block_blocked<T>(arena, xd, c0, o);
    // every wave's DINV stores complete, then publish
    publish_wait();
    __syncthreads();
    if (threadIdx.x == 0)
      __hip_atomic_store(P.progress, ((uint64_t)epoch << 32) | 1u, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  int local;
  const Prob& P = *probs[sched_member(rows, t - nd, local)];
  const int64_t ri = c0 + (int64_t)TILE * (local + 1);
  // the panel product runs in the storage type (as panel_rows)
  constexpr int PP = KP<T>::value;
  T* ck = (T*)&arena[0][0];
  T* bk = ck + TILE * PP;
  update_tile_to_lds<T, BF16, T>((const T*)P.L, P.Np, s0, j, ri, c0, (char*)arena, ck, PP, true);
  if (threadIdx.x == 0) wait_progress(P.progress, epoch, 1, P.info);
  __syncthreads();
  // B[p][n] = DINV_j(n, p), published in this launch: coherent loads
  const double* Dj = P.dinv + (size_t)j * TILE * TILE;
  {
    const int n = threadIdx.x & 63, pb = threadIdx.x >> 6;
    for (int q = 0; q < 16; ++q) {
      const int pp = pb + 4 * q;
      bk[pp * PP + n] = (T)ld_coherent(&Dj[n + pp * TILE]);
    }
  }
  __syncthreads();
  typename Acc<T>::type acc[2][2];
  zero_acc<T>(acc);
  mma64_kmajor<T>(ck, PP, bk, PP, acc);
  T* L = (T*)P.L;
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        L[(size_t)(ri + acc_row<T>(bm, r)) + (size_t)(c0 + acc_col(bn)) * P.Np] = acc[bm][bn][r];
}

// ---- column-form super-step in two launches ----------------------------------
// Phase A (k_superstep): the tiles (i, j) of the super-step's columns
// [s0, s1) with row block i < rlim (the super-step's diagonal block of
// tiles when rlim = s0 + G) in one launch, each tile one workgroup, in
// lookahead ticket order (StepSched).  A workgroup applies the left-looking
// update by the columns s0..j-1 as k_column does, fetching those columns'
// tiles -- written in this launch -- with coherent loads once their flags
// show them final; a diagonal workgroup then factors and inverts block j, a
// row workgroup waits for that and forms L_ij = C_ij DINV_j^T.  Each tile's
// flag is published after its stores complete.  The chain from one column to
// the next is tile (j+1, j) -> block j+1, so a column's factor overlaps the
// previous column's remaining rows, and the G - 1 launch gaps of the
// per-column form go away.  Phase B (k_panel_rows) takes the rows below,
// i >= rlim: one workgroup per row block, which forms its tiles (i, s0),
// (i, s0 + 1), ... in order from its own earlier tiles and phase A's
// (launch-ordered, so plain cached loads and no waits).  Same operations in
// the same order on every tile as k_column: the factor is bitwise the same.
//
// Member sync words (Prob::csync, zeroed when the descriptor is written with
// reset): flag (i, c) at csync[(c - s0) nb + i] holds the launch's epoch
// once tile (i, c) is final; csync[16 nb] counts the member's finished
// workgroups.  The member's last workgroup resets the count and clears the
// flags, so no flag outlives its launch (a replayed graph of the launch,
// whose epoch is the captured one, starts from cleared flags).
constexpr int MAXG = 16;
// Tickets run in groups k = s0 .. s1 (lookahead order): group k holds the
// diagonal workgroups D(k) of every member, then the rows i >= k + 1 of
// column k - 1 (B(k - 1)), then tile (k + 1, k) (the one block k + 1 needs
// from column k).  So the chain D(k) -> (k + 1, k) -> D(k + 1) is ticketed
// ahead of the bulk of each column, whose workgroups start once their
// diagonal block is (nearly) done instead of holding a slot while they
// wait; every dependency is on a lower ticket.
struct StepSched {
  int32_t count;             // members (sorted by size, descending)
  int32_t s0, s1;            // columns [s0, s1)
  int32_t rlim;              // row blocks [j, min(nb, rlim)) of each column j
  int32_t gbase[MAXG + 2];   // first ticket of group s0 + g
  int32_t nd[MAXG + 1];      // D(k): members with more than k block columns
  int32_t nB[MAXG + 1];      // members with min(nb, rlim) > k + 1 (B(k - 1) and tile (k + 1, k))
  int32_t nbk[MAXG + 1];     // workgroups of B(k - 1)
  int32_t pre[MAXB + 1];     // pre[m]: sum of min(nb, rlim) over the members before m
};

__device__ __forceinline__ uint32_t* tile_flag(const Prob& P, int nb, int s0, int i, int c) {
  return P.csync + (size_t)(c - s0) * nb + i;
}

// Wave 0 waits (bounded, one flag per lane) until the flags of tiles (i, c)
// and (j, c), c in [ca, cb), equal `epoch`; then the workgroup proceeds.
// All threads call it.
__device__ __forceinline__ void wait_tiles(const Prob& P, int nb, int s0, int i, int j, int ca,
                                           int cb, uint32_t epoch) {
  if (threadIdx.x < 64) {
    const int l = threadIdx.x;
    const int n = 2 * (cb - ca);
    const uint32_t* f = l < n ? tile_flag(P, nb, s0, (l & 1) ? j : i, ca + (l >> 1)) : nullptr;
    for (int spins = 0;; ++spins) {
      const bool ok =
          !f || __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch;
      if (__builtin_amdgcn_ballot_w64(!ok) == 0) break;
      if (spins >= SPIN_LIMIT) {
        if (l == 0) atomicOr(P.info, PGP_INFO_STALLED);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  __syncthreads();
}

template <typename T, int BF16>
__global__ __launch_bounds__(NTHREADS) void k_superstep(Prob* const* probs, StepSched ss,
                                                        uint32_t* ticket, uint32_t epoch) {
  __shared__ __attribute__((aligned(16))) char arena_b[col_arena_bytes<T, BF16>()];
  double (*arena)[TILE + 1] = (double (*)[TILE + 1])arena_b;
  double* xd = (double*)arena_b + TILE * (TILE + 1);
  __shared__ int s_val;
  const int t = take_ticket(ticket, &s_val);
  if ((unsigned)t >= gridDim.x) {  // counter not zero at launch: nothing to take
    flag_members_stalled(probs, ss.count);
    return;
  }
  // ticket -> (member mi, row block i, column j)
  int g = 0;
  while (t >= ss.gbase[g + 1]) ++g;
  const int s0 = ss.s0, k = s0 + g;
  int q = t - ss.gbase[g];
  int mi, i, j;
  if (q < ss.nd[g]) {  // D(k)
    mi = q;
    i = j = k;
  } else if ((q -= ss.nd[g]) < ss.nbk[g]) {
    // B(k - 1): member m has min(nb_m, rlim) - k - 1 rows, so the members
    // before m hold pre[m] - (k + 1) m of them
    int lo = 0, hi = ss.nB[g];
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (ss.pre[mid] - (k + 1) * mid <= q)
        lo = mid;
      else
        hi = mid;
    }
    mi = lo;
    i = k + 1 + (q - (ss.pre[lo] - (k + 1) * lo));
    j = k - 1;
  } else {  // tile (k + 1, k)
    mi = q - ss.nbk[g];
    i = k + 1;
    j = k;
  }
  const Prob& P = *probs[mi];
  const int nb = P.Np / TILE;
  const int64_t c0 = (int64_t)j * TILE, ri = (int64_t)i * TILE;
  // the columns before the last are normally final long before: one wait
  // for all of them, then column j - 1's tiles right before they are fetched
  if (j - 1 > s0) wait_tiles(P, nb, s0, i, j, s0, j - 1, epoch);
  auto last_col = [&](int c) {
    if (c == j - 1) wait_tiles(P, nb, s0, i, j, c, c + 1, epoch);
  };
  if (i == j) {
    update_tile_to_lds<T, BF16, double, true, true>((const T*)P.L, P.Np, s0, j, c0, c0,
                                                    (char*)arena, &arena[0][0], TILE + 1,
                                                    false, last_col);
    FactorOut o{P.L, P.Np, P.dinv + (size_t)j * TILE * TILE,
                P.logdiag ? P.logdiag + c0 : nullptr, P.diagval ? P.diagval + c0 : nullptr,
                P.info};
    // block j's rows wait only for DINV_j: publish before L_jj is stored
    factor_block_blocked<T>(arena, xd, c0, o, [&]() {
      publish_wait();
      __syncthreads();
      if (threadIdx.x == 0)
        __hip_atomic_store(tile_flag(P, nb, s0, j, j), epoch, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    });
  } else {
    constexpr int PP = KP<T>::value;
    T* ck = (T*)&arena[0][0];
    T* bk = ck + TILE * PP;
    update_tile_to_lds<T, BF16, T, false, true>((const T*)P.L, P.Np, s0, j, ri, c0, (char*)arena,
                                                ck, PP, true, last_col);
    wait_tiles(P, nb, s0, j, j, j, j + 1, epoch);  // block j factored, DINV_j published
    const double* Dj = P.dinv + (size_t)j * TILE * TILE;
    {
      const int n = threadIdx.x & 63, pb = threadIdx.x >> 6;
      for (int qq = 0; qq < 16; ++qq) {
        const int pp = pb + 4 * qq;
        bk[pp * PP + n] = (T)ld_coherent(&Dj[n + pp * TILE]);
      }
    }
    __syncthreads();
    typename Acc<T>::type acc[2][2];
    zero_acc<T>(acc);
    mma64_kmajor<T>(ck, PP, bk, PP, acc);
    T* L = (T*)P.L;
#pragma unroll
    for (int bm = 0; bm < 2; ++bm)
#pragma unroll
      for (int bn = 0; bn < 2; ++bn)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          st_coh(&L[(size_t)(ri + acc_row<T>(bm, r)) + (size_t)(c0 + acc_col(bn)) * P.Np],
                 (T)acc[bm][bn][r]);
  }
  // every wave's stores complete, then publish the tile
  publish_wait();
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_store(tile_flag(P, nb, s0, i, j), epoch, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    // the member's last workgroup in this launch clears its sync words
    const int e = min(ss.s1, nb), nbr = min(nb, ss.rlim);
    const uint32_t total = (uint32_t)((e - s0) * nbr - (s0 + e - 1) * (e - s0) / 2);
    uint32_t* done = P.csync + (size_t)MAXG * nb;
    s_val = (atomicAdd(done, 1u) + 1u == total) ? 1 : 0;
  }
  __syncthreads();
  if (s_val) {
    const int e = min(ss.s1, nb);
    for (int w = threadIdx.x; w < (e - s0) * nb; w += NTHREADS)
      __hip_atomic_store(P.csync + w, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x == 0)
      __hip_atomic_store(P.csync + (size_t)MAXG * nb, 0u, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Diagonal block k (k >= 1) after the trailing update of step k - 1 (separate
// mode): one workgroup per member.
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_factor_diag(Prob* const* probs, Sched sc, int k) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int64_t c0 = (int64_t)k * TILE;
  if (c0 >= P.Np) return;
  __shared__ double blk[TILE][TILE + 1];
  __shared__ double xd[4 * XD_BLK];
  const T* A = (const T*)P.L;
  const int m = threadIdx.x & 63, kb = threadIdx.x >> 6;
  for (int q = 0; q < 16; ++q) {
    const int c = kb + 4 * q;
    blk[m][c] = (double)A[(size_t)(c0 + m) + (size_t)(c0 + c) * P.Np];
  }
  __syncthreads();
  FactorOut o{P.L, P.Np, P.dinv + (size_t)k * TILE * TILE, P.logdiag ? P.logdiag + c0 : nullptr,
              P.diagval ? P.diagval + c0 : nullptr, P.info};
  factor_block_blocked<T>(blk, xd, c0, o);
}

// ---- raw (kernel-level ABI) variants ----------------------------------------
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_factor_first_raw(T* A, int64_t ld, double* dinv,
                                                               double* logdiag, int32_t* info) {
  FactorOut o{A, ld, dinv, logdiag, nullptr, info};
  factor_first<T>(A, ld, o);
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_panel_raw(T* A, int64_t N, int64_t ld,
                                                        const double* dinv, int k) {
  const int64_t c0 = (int64_t)k * TILE;
  const int64_t r0 = c0 + (int64_t)TILE * (blockIdx.x + 1);
  panel_rows<T>(A, ld, dinv + (size_t)k * TILE * TILE, c0, r0);
}

template <typename T, int BF16>
__global__ __launch_bounds__(NTHREADS) void k_syrk_raw(T* A, int64_t N, int64_t ld, int k,
                                                       double* dinv, double* logdiag,
                                                       int32_t* info) {
  const int64_t c0 = (int64_t)k * TILE, c1 = c0 + TILE;
  int ti, tj;
  lower_tile_c(blockIdx.x, ti, tj);
  const bool factor = (ti == 0 && tj == 0);
  FactorOut o{A, ld, dinv + (size_t)(k + 1) * TILE * TILE, logdiag ? logdiag + c1 : nullptr,
              nullptr, info};
  syrk_tile<T, BF16, 1>(A, ld, c0, c1 + (int64_t)TILE * ti, c1 + (int64_t)TILE * tj, factor, o);
}

// ---- inverse of every 64x64 diagonal block of a given factor (FP64) --------
// Used when a caller hands in a factor (kernel-level solves).  Lane i of each
// wave holds x_i of one column of inv(L_kk) during forward substitution.
template <typename T>
__device__ void invert_diag_block(const T* A, int64_t ld, int64_t c0, double* out) {
  __shared__ double Ls[TILE][TILE + 1];  // Ls[col][row]
  const int m = threadIdx.x & 63, kb = threadIdx.x >> 6;
  for (int q = 0; q < 16; ++q) {
    const int col = kb + 4 * q;
    Ls[col][m] = (m >= col) ? (double)A[(size_t)(c0 + m) + (size_t)(c0 + col) * ld] : 0.0;
  }
  __syncthreads();
  const int wv = threadIdx.x >> 6;
  const int i = threadIdx.x & 63;
  for (int cq = 0; cq < 16; ++cq) {
    const int c = wv * 16 + cq;
    double x = 0.0;
    double rhs = (i == c) ? 1.0 : 0.0;
    for (int j = c; j < 64; ++j) {
      if (i == j) x = rhs / Ls[j][j];
      const double xj = __shfl(x, j);
      if (i > j) rhs -= Ls[j][i] * xj;
    }
    out[(size_t)i + (size_t)c * TILE] = (i >= c) ? x : 0.0;
  }
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_dinv_raw(const T* A, int64_t ld, double* dinv) {
  invert_diag_block<T>(A, ld, (int64_t)blockIdx.x * TILE, dinv + (size_t)blockIdx.x * TILE * TILE);
}

// ---------------------------------------------------------------------------
// Schedule table (the autotuner's): per precision and size class, the
// super-step width G (rank-64G outer updates), COLUMN_MAX (super-steps whose
// first column has at most this many workgroups run in the column kernel;
// 0 disables), FUSE_MAX (update launches of at most this many tiles factor
// the next diagonal block inside the launch; 0 disables) and COLUMN_MERGE
// (how such a super-step is launched: 0 one k_column launch per column;
// 1 the diagonal block of tiles in one k_superstep launch, then the rows
// below in one k_panel_rows launch; 2 every tile in one k_superstep
// launch).
// Size class 0 holds the members with padded N <= TUNE_CLASS_NP, class 1 the
// larger ones; a batch is factored class by class, so a member's schedule --
// and with G its rounding -- depends only on its own size (a result is the
// same alone and in any mix, in the batcher or the single-problem path).
// COLUMN_MAX, FUSE_MAX and COLUMN_MERGE change only the launches, never the
// arithmetic.
// Initial values: PGP_SUPERSTEP / PGP_COLUMN_MAX / PGP_FUSE_MAX /
// PGP_COLUMN_MERGE from the environment, else 8 / 512 / 512 / 2.  Read at every factorisation, written
// rarely (pgp_tune_set): a mutex guards it.
// ---------------------------------------------------------------------------
constexpr int TUNE_CLASSES = 2;
constexpr int TUNE_CLASS_NP = 1024;

static int env_int(const char* name, int dflt) {
  const char* e = getenv(name);
  return e ? atoi(e) : dflt;
}

struct TuneTable {
  std::mutex mu;
  pgp_tune_entry_t e[3][TUNE_CLASSES];
  TuneTable() {
    const int G = std::min(16, std::max(1, env_int("PGP_SUPERSTEP", 8)));
    const int cm = env_int("PGP_COLUMN_MAX", 512), fm = env_int("PGP_FUSE_MAX", 512);
    const int mg = std::min(2, std::max(0, env_int("PGP_COLUMN_MERGE", 2)));
    for (int p = 0; p < 3; ++p)
      for (int c = 0; c < TUNE_CLASSES; ++c) e[p][c] = pgp_tune_entry_t{G, cm, fm, mg};
  }
};
static TuneTable& tune_table() {
  static TuneTable t;
  return t;
}

int tune_class_of(int np) { return np > TUNE_CLASS_NP ? 1 : 0; }

pgp_tune_entry_t tune_get(int prec, int cls) {
  TuneTable& t = tune_table();
  std::lock_guard<std::mutex> lock(t.mu);
  return t.e[prec][cls];
}

void tune_set(int prec, int cls, const pgp_tune_entry_t& v) {
  TuneTable& t = tune_table();
  std::lock_guard<std::mutex> lock(t.mu);
  t.e[prec][cls] = v;
}

static bool sorted_desc(const Batch& b) {
  for (int m = 1; m < b.count; ++m)
    if (b.np(m) > b.np(m - 1)) return false;
  return true;
}

template <typename T, int BF16>
static void column_launch(const Batch& b, int j, int s0, bool tpa, hipStream_t s) {
  constexpr double F3 = TILE * TILE * TILE / 3.0;
  constexpr double G64 = 2.0 * TILE * TILE * TILE;
  for (int off = 0; off < b.count; off += MAXB) {
    const int n = std::min(MAXB, b.count - off);
    Sched sc;
    sc.count = n;
    int tot = 0, nd = 0;
    double fl = 0.0;
    for (int m = 0; m < n; ++m) {
      const int nbm = b.np(off + m) / TILE;
      sc.pre[m] = tot;
      if (nbm > j) {
        ++nd;
        tot += nbm - j - 1;
        fl += F3 + (double)(nbm - j) * G64 * (j - s0) + (double)(nbm - j - 1) * G64;
      }
    }
    sc.pre[n] = tot;
    if (nd == 0) continue;
    const uint32_t ep = next_epoch();
    if (tpa) timer_begin(TC_PANEL, s);
    hipLaunchKernelGGL((k_column<T, BF16>), dim3(nd + tot), dim3(NTHREADS), 0, s, b.dprobs + off,
                       sc, nd, b.ticket, ep, j, s0);
    if (tpa) timer_end(TC_PANEL, s, fl, 0.0);
  }
}

// Phase B: row block i = s1 + local of a member with more than s1 block
// columns, columns [s0, s1): L_ij = (A_ij - sum_{c in [s0, j)} L_ic L_jc^T)
// DINV_j^T, j in order.  L_jc and DINV_j come from phase A (the previous
// launch), L_ic from this workgroup's earlier iterations.
template <typename T, int BF16>
struct RowsArena {
  static constexpr int PP = KP<T>::value;
  static constexpr int OPND = BF16 ? 72 * 2 : (MmajorF32<T, BF16>::value ? 68 * 4 : PP * (int)sizeof(T));
  static constexpr int A = 2 * TILE * OPND;                   // update operands
  static constexpr int B = 2 * TILE * PP * (int)sizeof(T);  // C_ij and DINV_j
  static constexpr int bytes = A > B ? A : B;
};

template <typename T, int BF16>
__global__ __launch_bounds__(NTHREADS) void k_panel_rows(Prob* const* probs, Sched sc, int s0,
                                                         int s1) {
  constexpr int PP = KP<T>::value;
  __shared__ __attribute__((aligned(16))) char arena[RowsArena<T, BF16>::bytes];
  int local;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, local)];
  const int i = s1 + local;
  if (i >= P.Np / TILE) return;
  const int64_t ri = (int64_t)i * TILE;
  T* ck = (T*)arena;
  T* bk = ck + TILE * PP;
  T* L = (T*)P.L;
  const int n = threadIdx.x & 63, pb = threadIdx.x >> 6;
  for (int j = s0; j < s1; ++j) {
    const int64_t c0 = (int64_t)j * TILE;
    update_tile_to_lds<T, BF16, T>((const T*)P.L, P.Np, s0, j, ri, c0, arena, ck, PP, true);
    const double* Dj = P.dinv + (size_t)j * TILE * TILE;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int pp = pb + 4 * q;
      bk[pp * PP + n] = (T)Dj[n + pp * TILE];
    }
    __syncthreads();
    typename Acc<T>::type acc[2][2];
    zero_acc<T>(acc);
    mma64_kmajor<T>(ck, PP, bk, PP, acc);
#pragma unroll
    for (int bm = 0; bm < 2; ++bm)
#pragma unroll
      for (int bn = 0; bn < 2; ++bn)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          L[(size_t)(ri + acc_row<T>(bm, r)) + (size_t)(c0 + acc_col(bn)) * P.Np] =
              acc[bm][bn][r];
    // the tile is an operand of this row's next columns; LDS is free again
    publish_wait();
    __syncthreads();
  }
}

template <typename T, int BF16>
static void rows_launch(const Batch& b, int s0, int s1, bool tpa, hipStream_t s) {
  constexpr double G64 = 2.0 * TILE * TILE * TILE;
  sched_launch(
      b, [&](int m) { return std::max(0, b.np(m) / TILE - s1); },
      [&](Prob* const* dp, const Sched& sc, int n) {
        if (tpa) timer_begin(TC_PANEL, s);
        hipLaunchKernelGGL((k_panel_rows<T, BF16>), dim3(n), dim3(NTHREADS), 0, s, dp, sc, s0, s1);
        if (tpa) {
          const int off = (int)(dp - b.dprobs);
          double fl = 0.0;
          const double per_row = G64 * (double)(s1 - s0) * (s1 - s0 + 1) / 2.0;
          for (int m = off; m < off + sc.count; ++m)
            fl += std::max(0, b.np(m) / TILE - s1) * per_row;
          timer_end(TC_PANEL, s, fl, 0.0);
        }
      });
}

template <typename T, int BF16>
static void superstep_launch(const Batch& b, int s0, int s1, int rlim, bool tpa, hipStream_t s) {
  constexpr double F3 = TILE * TILE * TILE / 3.0;
  constexpr double G64 = 2.0 * TILE * TILE * TILE;
  for (int off = 0; off < b.count; off += MAXB) {
    const int n = std::min(MAXB, b.count - off);
    StepSched ss;
    ss.count = n;
    ss.s0 = s0;
    ss.s1 = s1;
    ss.rlim = rlim;
    ss.pre[0] = 0;
    for (int m = 0; m < n; ++m) ss.pre[m + 1] = ss.pre[m] + std::min(b.np(off + m) / TILE, rlim);
    double fl = 0.0;
    for (int m = 0; m < n; ++m) {
      const int nbm = b.np(off + m) / TILE, nbr = std::min(nbm, rlim);
      for (int j = s0; j < std::min(s1, nbm); ++j)
        fl += F3 + (double)(nbr - j) * G64 * (j - s0) + (double)(nbr - j - 1) * G64;
    }
    const int ncols = s1 - s0;
    ss.gbase[0] = 0;
    for (int g = 0; g <= ncols; ++g) {
      const int k = s0 + g;
      int nd = 0, nB = 0, nbk = 0;
      for (int m = 0; m < n; ++m) {
        const int nbm = b.np(off + m) / TILE, nbr = std::min(nbm, rlim);
        if (g < ncols && nbm > k) ++nd;
        if (nbr > k + 1) {
          ++nB;
          if (g >= 1) nbk += nbr - k - 1;
        }
      }
      ss.nd[g] = nd;
      ss.nB[g] = nB;
      ss.nbk[g] = nbk;
      ss.gbase[g + 1] = ss.gbase[g] + nd + nbk + (g < ncols ? nB : 0);
    }
    const int tot = ss.gbase[ncols + 1];
    if (tot == 0) continue;
    const uint32_t ep = next_epoch();
    const bool tsu = timer_on(TC_SUPER);
    double by = 0.0;
    if (tsu) {
      // algorithmic bytes: every tile of the launch read and written once in
      // the storage type, and each diagonal block's FP64 inverse written once
      constexpr double TB = (double)TILE * TILE;
      for (int m = 0; m < n; ++m) {
        const int nbm = b.np(off + m) / TILE, nbr = std::min(nbm, rlim);
        for (int j = s0; j < std::min(s1, nbm); ++j)
          by += (double)(nbr - j) * 2.0 * TB * sizeof(T) + TB * 8.0;
      }
      timer_begin(TC_SUPER, s);
    }
    hipLaunchKernelGGL((k_superstep<T, BF16>), dim3(tot), dim3(NTHREADS), 0, s, b.dprobs + off, ss,
                       b.ticket, ep);
    if (tsu) timer_end(TC_SUPER, s, fl, by);
  }
}

// ---------------------------------------------------------------------------
// Per block column k (every member at once), in super-steps of G columns:
//   separate kernels: U_k left-looking update of column k by the super-step's
//     earlier panels; F_k factor of block k (inside U_k when that launch is
//     small); P_k panel; at the super-step's last column O, the outer update
//     of everything right of it (factoring its first block when small);
//   column kernel: one k_column launch per column (the same update, factor
//     and panel), then O at the super-step's end.
// Both forms apply the same operations in the same order to every tile, so a
// member's factor does not depend on which form its batch used.
template <typename T, int BF16>
static void potrf_steps(const Batch& b, hipStream_t s, const pgp_tune_entry_t& te) {
  const int nb = b.maxNp / TILE;
  const int G = te.superstep;
  const int column_max = te.column_max, fuse_max = te.fuse_max;
  const bool tsy = timer_on(TC_SYRK), tpa = timer_on(TC_PANEL);
  constexpr double F3 = TILE * TILE * TILE / 3.0;  // factor of one 64x64 block
  constexpr double G64 = 2.0 * TILE * TILE * TILE;  // one 64x64x64 product
  // algorithmic flops of the members [off, off + n) of a launch
  auto flops = [&](const Sched& sc, Prob* const* dp, auto per_member) {
    const int off = (int)(dp - b.dprobs);
    double f = 0.0;
    for (int m = off; m < off + sc.count; ++m) f += per_member(b.np(m) / TILE);
    return f;
  };
  const bool colok = column_max > 0 && b.ticket && sorted_desc(b);
  // workgroups of the column launch at the first column of a super-step;
  // non-increasing in s0, so once a super-step uses the column kernel all
  // later ones do
  auto col_style = [&](int s0) {
    if (!colok) return false;
    long long w = 0;
    for (int m = 0; m < b.count; ++m) w += std::max(0, b.np(m) / TILE - s0);
    return w <= column_max;
  };
  auto outer_launch = [&](int s0, int s_end, bool fuse) {
    auto outer_tiles = [&](int m) {
      const int t = std::max(0, b.np(m) / TILE - s_end);
      return t * (t + 1) / 2;
    };
    if constexpr (BF16 != 0) {
      // the BF16 copy of the panel rows below the super-step, which the
      // outer update reads instead of converting at every stash
      sched_launch(
          b, [&](int m) { return std::max(0, b.np(m) / TILE - s_end); },
          [&](Prob* const* dp, const Sched& sc, int n) {
            hipLaunchKernelGGL(k_panel_bf16, dim3(n), dim3(NTHREADS), 0, s, dp, sc, s0, s_end);
          });
    }
    sched_launch(b, outer_tiles, [&](Prob* const* dp, const Sched& sc, int n) {
      if (tsy) timer_begin(TC_SYRK, s);
      if (fuse)
        hipLaunchKernelGGL((k_syrk_outer<T, BF16, 1>), dim3(n), dim3(NTHREADS), 0, s, dp, sc, s0,
                           s_end);
      else
        hipLaunchKernelGGL((k_syrk_outer<T, BF16, 0>), dim3(n), dim3(NTHREADS), 0, s, dp, sc, s0,
                           s_end);
      if (tsy) {
        // outer update: an n x n lower triangle (n = 64 T) with rank K = 64 G
        // counts n (n + 1) K flops; bytes: every lower tile read and written
        // once, each trailing row block's G panel blocks read once
        const double es = sizeof(T);
        timer_end(TC_SYRK, s, flops(sc, dp,
                                    [&](int nbm) {
                                      const double nn = (double)std::max(0, nbm - s_end) * TILE;
                                      return nn * (nn + 1) * (double)(TILE * G);
                                    }),
                  flops(sc, dp, [&](int nbm) {
                    // (BF16 factors read the panel from its BF16 copy)
                    const double t = std::max(0, nbm - s_end);
                    return t * (t + 1) * TILE * TILE * es +
                           t * TILE * TILE * G * (BF16 ? 2.0 : es);
                  }));
      }
    });
  };
  auto total_units = [&](auto units) {
    int t = 0;
    for (int m = 0; m < b.count; ++m) t += units(m);
    return t;
  };
  if (!col_style(0)) {
    if (tpa) timer_begin(TC_PANEL, s);
    hipLaunchKernelGGL(k_factor_first<T>, dim3(1, b.count), dim3(NTHREADS), 0, s, b.dprobs);
    if (tpa) timer_end(TC_PANEL, s, b.count * F3, 0.0);
  }
  for (int k = 0; k < nb; ++k) {
    const int s0 = k / G * G, s_end = s0 + G;
    const bool outer = (k + 1 == s_end);
    if (col_style(s0)) {
      if (te.column_merge == 2) {
        // every row block of the super-step's columns in the one launch
        if (k == s0) superstep_launch<T, BF16>(b, s0, std::min(s_end, nb), nb, tpa, s);
      } else if (te.column_merge == 1) {
        if (k == s0) {
          superstep_launch<T, BF16>(b, s0, std::min(s_end, nb), s_end, tpa, s);
          if (s_end < nb) rows_launch<T, BF16>(b, s0, s_end, tpa, s);
        }
      } else {
        column_launch<T, BF16>(b, k, s0, tpa, s);
      }
      if (outer && s_end < nb) outer_launch(s0, s_end, false);
      continue;
    }
    // U_k (k > s0): left-looking update of column k, factoring block k inside
    // the launch when it is small, else F_k after it.  Block s0 of a
    // super-step was factored at the end of the previous one (or by F_0).
    if (k > s0) {
      auto tiles = [&](int m) { return std::max(0, b.np(m) / TILE - k); };
      const bool fused = total_units(tiles) <= fuse_max;
      sched_launch(b, tiles, [&](Prob* const* dp, const Sched& sc, int n) {
        if (tpa) timer_begin(TC_PANEL, s);
        if (fused)
          hipLaunchKernelGGL((k_col_update<T, BF16, 1>), dim3(n), dim3(NTHREADS), 0, s, dp, sc, k,
                             s0);
        else
          hipLaunchKernelGGL((k_col_update<T, BF16, 0>), dim3(n), dim3(NTHREADS), 0, s, dp, sc, k,
                             s0);
        if (tpa)
          timer_end(TC_PANEL, s,
                    flops(sc, dp, [&](int nbm) {
                      return nbm > k ? (double)(nbm - k) * G64 * (k - s0) + (fused ? F3 : 0.0)
                                     : 0.0;
                    }),
                    0.0);
      });
      if (!fused)
        sched_launch(
            b, [&](int m) { return b.np(m) / TILE > k ? 1 : 0; },
            [&](Prob* const* dp, const Sched& sc, int n) {
              if (tpa) timer_begin(TC_PANEL, s);
              hipLaunchKernelGGL(k_factor_diag<T>, dim3(n), dim3(NTHREADS), 0, s, dp, sc, k);
              if (tpa) timer_end(TC_PANEL, s, n * F3, 0.0);
            });
    }
    if (k + 1 >= nb) break;
    sched_launch(
        b, [&](int m) { return std::max(0, b.np(m) / TILE - k - 1); },
        [&](Prob* const* dp, const Sched& sc, int n) {
          if (tpa) timer_begin(TC_PANEL, s);
          hipLaunchKernelGGL(k_panel<T>, dim3(n), dim3(NTHREADS), 0, s, dp, sc, k);
          if (tpa)
            timer_end(TC_PANEL, s,
                      flops(sc, dp, [&](int nbm) { return std::max(0, nbm - k - 1) * G64; }), 0.0);
        });
    if (outer) {
      // O closes the super-step; block s_end is factored inside it when the
      // launch is small, else by F_s_end -- unless the next super-step uses
      // the column kernel, which factors it itself
      const bool next_col = col_style(s_end);
      const bool fused = !next_col && total_units([&](int m) {
                                        const int t = std::max(0, b.np(m) / TILE - s_end);
                                        return t * (t + 1) / 2;
                                      }) <= fuse_max;
      outer_launch(s0, s_end, fused);
      if (!fused && !next_col)
        sched_launch(
            b, [&](int m) { return b.np(m) / TILE > s_end ? 1 : 0; },
            [&](Prob* const* dp, const Sched& sc, int n) {
              if (tpa) timer_begin(TC_PANEL, s);
              hipLaunchKernelGGL(k_factor_diag<T>, dim3(n), dim3(NTHREADS), 0, s, dp, sc, s_end);
              if (tpa) timer_end(TC_PANEL, s, n * F3, 0.0);
            });
    }
  }
}

static void potrf_class(const Batch& b, int prec, hipStream_t s) {
  if (b.count <= 0) return;
  const pgp_tune_entry_t te = tune_get(prec, tune_class_of(b.maxNp));
  if (prec == PGP_PREC_FP64)
    potrf_steps<double, 0>(b, s, te);
  else if (prec == PGP_PREC_FP32)
    potrf_steps<float, 0>(b, s, te);
  else
    potrf_steps<float, 1>(b, s, te);
}

// The members of a batch are sorted by size (largest first), so the large
// size class is a prefix: each class is factored with its own schedule.
void launch_potrf(const Batch& b, int prec, hipStream_t s) {
  int k = 0;
  while (k < b.count && tune_class_of(b.np(k)) == 1) ++k;
  if (k == 0 || k == b.count) {
    potrf_class(b, prec, s);
    return;
  }
  Batch big = b, small = b;
  big.count = k;
  small.dprobs = b.dprobs + k;
  small.count = b.count - k;
  small.maxNp = b.np(k);
  if (b.hnp) small.hnp = b.hnp + k;
  if (b.hm) small.hm = b.hm + k;
  potrf_class(big, prec, s);
  potrf_class(small, prec, s);
}

// The factorisation leaves DINV in place; kept as a no-op entry for callers
// that re-factor through launch_potrf.
void launch_dinv(const Batch& b, int prec, hipStream_t s) {
  (void)b;
  (void)prec;
  (void)s;
}

template <typename T, int BF16>
static void potrf_raw_steps(T* A, int64_t N, int64_t ld, double* dinv, double* logdiag,
                            int32_t* info, hipStream_t s) {
  const int nb = (int)(N / TILE);
  hipLaunchKernelGGL(k_factor_first_raw<T>, dim3(1), dim3(NTHREADS), 0, s, A, ld, dinv, logdiag,
                     info);
  for (int k = 0; k + 1 < nb; ++k) {
    const int below = nb - k - 1;
    hipLaunchKernelGGL(k_panel_raw<T>, dim3(below), dim3(NTHREADS), 0, s, A, N, ld,
                       (const double*)dinv, k);
    hipLaunchKernelGGL((k_syrk_raw<T, BF16>), dim3(below * (below + 1) / 2), dim3(NTHREADS), 0, s,
                       A, N, ld, k, dinv, logdiag, info);
  }
}

int launch_potrf_raw(int prec, void* A, int64_t N, int64_t ld, double* dinv, double* logdiag,
                     int32_t* info, hipStream_t s) {
  if (prec == PGP_PREC_FP64)
    potrf_raw_steps<double, 0>((double*)A, N, ld, dinv, logdiag, info, s);
  else if (prec == PGP_PREC_FP32)
    potrf_raw_steps<float, 0>((float*)A, N, ld, dinv, logdiag, info, s);
  else
    potrf_raw_steps<float, 1>((float*)A, N, ld, dinv, logdiag, info, s);
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
