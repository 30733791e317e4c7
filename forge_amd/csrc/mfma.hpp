// The 16 x 16 x 4 matrix-core instructions of gfx950 per scalar type, shared
// by the GEMM kernels (gemm.hip) and the two-stage reduction (twostage.hip).
#pragma once
#include "common.hpp"

namespace forge {

// Workgroups are dispatched round-robin over the 8 XCDs (linear id % 8), each
// with its own L2.  xcd_remap(pid, nb) maps the linear id to a logical one so
// that every XCD works on a contiguous range of the nb logical blocks, and
// neighbouring blocks that read the same operands share one L2.
__device__ __forceinline__ i64 xcd_remap(i64 pid, i64 nb) {
  const i64 q = nb / 8;
  if (pid >= q * 8) return pid;
  return (pid % 8) * q + pid / 8;
}

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

// A complex 16 x 16 product accumulated on the real instructions, operands
// (ar + i ai)(br + i bi) one value per lane as above.  M3 = false: four real
// products into re and im.  M3 = true: Gauss's three, P1 = ar br, P2 = ai bi,
// P3 = (ar + ai)(br + bi), with re = P1 - P2 and im = P3 - P1 - P2 formed at
// the end -- a quarter fewer matrix-core operations for an imaginary part
// whose rounding is bounded by (|ar| + |ai|)(|br| + |bi|) instead of
// |ar||bi| + |ai||br|.
//   begin(cr, ci)  start a sum that ends added to (cr, ci): the four-product
//                  form accumulates into them directly
//   end(cr, ci)    (cr, ci) += the sum
template <typename S, bool M3> struct CAcc;
template <typename S> struct CAcc<S, false> {
  typedef Mfma<S> MF;
  typename MF::acc r, m;
  __device__ __forceinline__ void zero() { r = typename MF::acc{}; m = typename MF::acc{}; }
  __device__ __forceinline__ void begin(const typename MF::acc& cr, const typename MF::acc& ci) {
    r = cr;
    m = ci;
  }
  __device__ __forceinline__ void mma(S ar, S ai, S br, S bi) {
    r = MF::mma(ar, br, r);
    m = MF::mma(ar, bi, m);
    r = MF::mma(-ai, bi, r);
    m = MF::mma(ai, br, m);
  }
  __device__ __forceinline__ S re(int i) const { return r[i]; }
  __device__ __forceinline__ S im(int i) const { return m[i]; }
  __device__ __forceinline__ void end(typename MF::acc& cr, typename MF::acc& ci) const {
    cr = r;
    ci = m;
  }
};
template <typename S> struct CAcc<S, true> {
  typedef Mfma<S> MF;
  typename MF::acc p1, p2, p3;
  __device__ __forceinline__ void zero() {
    p1 = typename MF::acc{};
    p2 = typename MF::acc{};
    p3 = typename MF::acc{};
  }
  __device__ __forceinline__ void begin(const typename MF::acc&, const typename MF::acc&) { zero(); }
  __device__ __forceinline__ void mma(S ar, S ai, S br, S bi) {
    p1 = MF::mma(ar, br, p1);
    p2 = MF::mma(ai, bi, p2);
    p3 = MF::mma(ar + ai, br + bi, p3);
  }
  __device__ __forceinline__ S re(int i) const { return p1[i] - p2[i]; }
  __device__ __forceinline__ S im(int i) const { return p3[i] - p1[i] - p2[i]; }
  __device__ __forceinline__ void end(typename MF::acc& cr, typename MF::acc& ci) const {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      cr[i] += re(i);
      ci[i] += im(i);
    }
  }
};

}  // namespace forge
