// 64 x 64 x 64 tile products on CDNA MFMA units.
//
// A workgroup of 256 threads (4 waves) computes C(64x64) += A(64x64) B(64x64);
// wave w owns the 32 x 32 quadrant (rows 32*(w>>1), cols 32*(w&1)) as 2 x 2
// blocks of 16 x 16.  Operands come from LDS in "k-major" layout
// As[k * pitch + m] = A(m, k), Bs[k * pitch + n] = B(k, n), which is how a
// column-major global panel lands in LDS without a transpose.  The bf16 path
// uses "m-major" layout (8 consecutive k per 16-byte read).
//
// Lane layout of the 16x16xK MFMA family (lane l = threadIdx.x & 63):
//   A operand: row (l & 15), k (l >> 4) [x8 consecutive k for bf16 x32]
//   B operand: col (l & 15), k (l >> 4) [x8 consecutive k for bf16 x32]
//   D (f32, bf16 inputs): row 4*(l >> 4) + r (r = 0..3), col (l & 15)
//   D (f64 16x16x4):      row (l >> 4) + 4*r  (r = 0..3), col (l & 15)
// (the f64 row interleave was measured on gfx950 with a lane-layout probe).
#pragma once
#include <hip/hip_runtime.h>

namespace pgp {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// Pitch (elements) of a k-major 64-column LDS operand.  For 4-byte types the
// four k-groups of a 16x16x4 MFMA operand read (lanes 16q.., rows r..r+15,
// k = kk + q) then fall on disjoint LDS banks (pitch = 16 mod 32 banks);
// 8-byte types already do with 65.
template <typename T>
struct KP {
  static constexpr int value = sizeof(T) == 4 ? 80 : 65;
};

template <typename T>
struct Acc;
template <>
struct Acc<double> {
  typedef doublex4 type;
};
template <>
struct Acc<float> {
  typedef floatx4 type;
};

__device__ __forceinline__ doublex4 mfma16(double a, double b, doublex4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ floatx4 mfma16(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Round-to-nearest-even float -> bf16 bits.
__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

template <typename T>
__device__ __forceinline__ void zero_acc(typename Acc<T>::type (&acc)[2][2]) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = (T)0;
}

// acc += A B over K = 64, k-major LDS operands (double / float).
template <typename T>
__device__ __forceinline__ void mma64_kmajor(const T* As, int pa, const T* Bs, int pb,
                                             typename Acc<T>::type (&acc)[2][2]) {
  const int l = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int wr = 32 * (w >> 1), wc = 32 * (w & 1);
  const int lr = l & 15, lk = l >> 4;
#pragma unroll 4
  for (int kk = 0; kk < 64; kk += 4) {
    const T a0 = As[(kk + lk) * pa + wr + lr];
    const T a1 = As[(kk + lk) * pa + wr + 16 + lr];
    const T b0 = Bs[(kk + lk) * pb + wc + lr];
    const T b1 = Bs[(kk + lk) * pb + wc + 16 + lr];
    acc[0][0] = mfma16(a0, b0, acc[0][0]);
    acc[0][1] = mfma16(a0, b1, acc[0][1]);
    acc[1][0] = mfma16(a1, b0, acc[1][0]);
    acc[1][1] = mfma16(a1, b1, acc[1][1]);
  }
}

// mma64_kmajor over K = 32 (k-major operands of 32 rows): the first or the
// second half of mma64_kmajor's k steps, in the same order.
template <typename T>
__device__ __forceinline__ void mma64_kmajor_k32(const T* As, int pa, const T* Bs, int pb,
                                                 typename Acc<T>::type (&acc)[2][2]) {
  const int l = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int wr = 32 * (w >> 1), wc = 32 * (w & 1);
  const int lr = l & 15, lk = l >> 4;
#pragma unroll 4
  for (int kk = 0; kk < 32; kk += 4) {
    const T a0 = As[(kk + lk) * pa + wr + lr];
    const T a1 = As[(kk + lk) * pa + wr + 16 + lr];
    const T b0 = Bs[(kk + lk) * pb + wc + lr];
    const T b1 = Bs[(kk + lk) * pb + wc + 16 + lr];
    acc[0][0] = mfma16(a0, b0, acc[0][0]);
    acc[0][1] = mfma16(a0, b1, acc[0][1]);
    acc[1][0] = mfma16(a1, b0, acc[1][0]);
    acc[1][1] = mfma16(a1, b1, acc[1][1]);
  }
}

// acc += A B^T over K = 64 for float, m-major LDS operands As[m * pa + k],
// Bs[n * pb + k] (pa, pb multiples of 4; with pa = pb = 4 mod 32 the 16-byte
// reads of 8 consecutive lanes cover the 32 banks once; the counters still
// show a few conflict cycles).  MFMA step s gives lane group q = l >> 4 the
// k index 16 q + s, so a lane's operands for a half of the chunk are 8
// consecutive floats: two 16-byte LDS reads per operand and half, against
// one 4-byte read per operand and step in the k-major form.  The k order
// inside each MFMA differs from mma64_kmajor's, so every kernel whose
// results must agree bit for bit uses the same one of the two.
__device__ __forceinline__ void mma64_f32_mmajor(const float* As, int pa, const float* Bs, int pb,
                                                 floatx4 (&acc)[2][2]) {
  const int l = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int wr = 32 * (w >> 1), wc = 32 * (w & 1);
  const int lr = l & 15, k0 = 16 * (l >> 4);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    floatx4 a0[2], a1[2], b0[2], b1[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int k = k0 + 8 * h + 4 * j;
      a0[j] = *(const floatx4*)&As[(wr + lr) * pa + k];
      a1[j] = *(const floatx4*)&As[(wr + 16 + lr) * pa + k];
      b0[j] = *(const floatx4*)&Bs[(wc + lr) * pb + k];
      b1[j] = *(const floatx4*)&Bs[(wc + 16 + lr) * pb + k];
    }
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      acc[0][0] = mfma16(a0[s >> 2][s & 3], b0[s >> 2][s & 3], acc[0][0]);
      acc[0][1] = mfma16(a0[s >> 2][s & 3], b1[s >> 2][s & 3], acc[0][1]);
      acc[1][0] = mfma16(a1[s >> 2][s & 3], b0[s >> 2][s & 3], acc[1][0]);
      acc[1][1] = mfma16(a1[s >> 2][s & 3], b1[s >> 2][s & 3], acc[1][1]);
    }
  }
}

// acc += A B^T-style product for bf16: As[m * pa + k], Bs[n * pb + k]
// (m-major / n-major bf16 bits), K = 64 as two x32 steps.
__device__ __forceinline__ void mma64_bf16(const unsigned short* As, int pa,
                                           const unsigned short* Bs, int pb,
                                           floatx4 (&acc)[2][2]) {
  const int l = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int wr = 32 * (w >> 1), wc = 32 * (w & 1);
  const int lr = l & 15, lk = 8 * (l >> 4);
#pragma unroll
  for (int kk = 0; kk < 64; kk += 32) {
    const u16x8 a0 = *(const u16x8*)&As[(wr + lr) * pa + kk + lk];
    const u16x8 a1 = *(const u16x8*)&As[(wr + 16 + lr) * pa + kk + lk];
    const u16x8 b0 = *(const u16x8*)&Bs[(wc + lr) * pb + kk + lk];
    const u16x8 b1 = *(const u16x8*)&Bs[(wc + 16 + lr) * pb + kk + lk];
    const bf16x8 A0 = __builtin_bit_cast(bf16x8, a0);
    const bf16x8 A1 = __builtin_bit_cast(bf16x8, a1);
    const bf16x8 B0 = __builtin_bit_cast(bf16x8, b0);
    const bf16x8 B1 = __builtin_bit_cast(bf16x8, b1);
    acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B1, acc[1][1], 0, 0, 0);
  }
}

// Row / column of accumulator element (bm, bn, r) relative to the 64x64 tile.
template <typename T>
__device__ __forceinline__ int acc_row(int bm, int r) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  return 32 * (w >> 1) + 16 * bm + 4 * (l >> 4) + r;
}
template <>
__device__ __forceinline__ int acc_row<double>(int bm, int r) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  return 32 * (w >> 1) + 16 * bm + (l >> 4) + 4 * r;
}
__device__ __forceinline__ int acc_col(int bn) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  return 32 * (w & 1) + 16 * bn + (l & 15);
}

// Load a 64x64 column-major global block (rows r0.., cols c0.., ld) into
// k-major LDS dst[k * pitch + m] = G(r0 + m, c0 + k), converting to TL.
template <typename TG, typename TL>
__device__ __forceinline__ void load_block_kmajor(TL* dst, int pitch, const TG* G, int64_t ld,
                                                  int64_t r0, int64_t c0) {
  const int m = threadIdx.x & 63;
  const int kb = threadIdx.x >> 6;
#pragma unroll 4
  for (int q = 0; q < 16; ++q) {
    const int k = kb + 4 * q;
    dst[k * pitch + m] = (TL)G[(size_t)(r0 + m) + (size_t)(c0 + k) * ld];
  }
}

// Load a 64x64 column-major block into m-major bf16 LDS dst[m * pitch + k].
template <typename TG>
__device__ __forceinline__ void load_block_bf16_mmajor(unsigned short* dst, int pitch, const TG* G,
                                                       int64_t ld, int64_t r0, int64_t c0) {
  const int m = threadIdx.x & 63;
  const int kb = threadIdx.x >> 6;
#pragma unroll 4
  for (int q = 0; q < 16; ++q) {
    const int k = kb + 4 * q;
    dst[m * pitch + k] = f2bf((float)G[(size_t)(r0 + m) + (size_t)(c0 + k) * ld]);
  }
}

}  // namespace pgp
