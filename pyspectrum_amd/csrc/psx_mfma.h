// Operand helpers of the split-BF16 MFMA contractions (gfx950).
#pragma once
#include "psx_common.h"

namespace psx {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
typedef double d4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef unsigned short us8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

typedef float f2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
typedef unsigned u2v __attribute__((ext_vector_type(2)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));

// Two floats -> packed bf16 pair (round to nearest even; v_cvt_pk_bf16_f32).
__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
  f2v x;
  x[0] = a; x[1] = b;
  return __builtin_bit_cast(unsigned, __builtin_convertvector(x, bf2v));
}

// Residuals of a packed bf16 pair hb = (bf16(x0), bf16(x1)): x0 - hi0 and
// x1 - hi1, each one v_dot2_f32_bf16 (hb . (-1, 0) + x0, hb . (0, -1) + x1).
// The difference of a float and its nearest bf16 is exactly representable,
// so the single rounding of the dot product leaves it exact.
// The (-1, 0) and (0, -1) pairs are passed in registers: an inline constant
// for a packed-bf16 operand is not the same bit pattern on every path.
__device__ __forceinline__ void bf16_residuals(unsigned hb, float x0, float x1, unsigned neg_lo,
                                               unsigned neg_hi, float& r0, float& r1) {
  const bf2v h = __builtin_bit_cast(bf2v, hb);
  r0 = __builtin_amdgcn_fdot2_f32_bf16(h, __builtin_bit_cast(bf2v, neg_lo), x0, false);
  r1 = __builtin_amdgcn_fdot2_f32_bf16(h, __builtin_bit_cast(bf2v, neg_hi), x1, false);
}

// A value the compiler must keep in a register (no constant folding).
__device__ __forceinline__ unsigned in_register(unsigned v) {
  asm volatile("" : "+v"(v));
  return v;
}

// Products p = a*b of 8 point pairs, split as p = hi + lo: hi = bf16(p), lo =
// bf16(p - hi).  Per value: a multiply, half a pack, one dot2 for the
// residual, half a pack.
__device__ __forceinline__ void product_split_bf16(const f4v& a0, const f4v& a1, const f4v& b0,
                                                   const f4v& b1, unsigned neg_lo, unsigned neg_hi,
                                                   bf8v& hi, bf8v& lo) {
  float x[8];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    x[e] = a0[e] * b0[e];
    x[4 + e] = a1[e] * b1[e];
  }
  u4v h, l;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const unsigned hb = pack_bf16(x[2 * p], x[2 * p + 1]);
    float r0, r1;
    bf16_residuals(hb, x[2 * p], x[2 * p + 1], neg_lo, neg_hi, r0, r1);
    h[p] = hb;
    l[p] = pack_bf16(r0, r1);
  }
  hi = __builtin_bit_cast(bf8v, h);
  lo = __builtin_bit_cast(bf8v, l);
}

}  // namespace psx
