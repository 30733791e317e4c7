// The 16 x 16 x 4 matrix-core instructions of gfx950 per scalar type, shared
// by the GEMM kernels (gemm.hip) and the two-stage reduction (twostage.hip).
#pragma once
#include "common.hpp"

namespace forge {

typedef double gf64x4 __attribute__((ext_vector_type(4)));
typedef float gf32x4 __attribute__((ext_vector_type(4)));

// Inputs one value per lane: A(row lane % 16, k lane / 16), B(k lane / 16,
// column lane % 16).  Output four values per lane in column lane % 16 at
// rows row(lane, i), i = 0..3 (the FP64 and FP32 instructions place them
// differently; both layouts were checked on the device).
template <typename S> struct Mfma;
template <> struct Mfma<double> {
  typedef gf64x4 acc;
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int i) { return (lane >> 4) + 4 * i; }
};
template <> struct Mfma<float> {
  typedef gf32x4 acc;
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int i) { return 4 * (lane >> 4) + i; }
};

}  // namespace forge
