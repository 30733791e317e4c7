// Shared device/host helpers of the MI355X Hartree-Fock engine (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include <string>

#include "../../include/forge_hf.h"

namespace forge {

typedef int64_t i64;

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
#define FORGE_HIP(call)                                                              \
  do {                                                                               \
    hipError_t _e = (call);                                                          \
    if (_e != hipSuccess) {                                                          \
      ::forge::set_error(std::string(#call) + ": " + hipGetErrorString(_e) + " (" + \
                         __FILE__ + ":" + std::to_string(__LINE__) + ")");           \
      return -1;                                                                     \
    }                                                                                \
  } while (0)
#define FORGE_LAUNCH() FORGE_HIP(hipGetLastError())

// ---------------------------------------------------------------- complex
struct z2 {
  double x, y;
};
__host__ __device__ inline z2 mk(double a, double b) { return z2{a, b}; }
__host__ __device__ inline z2 operator+(z2 a, z2 b) { return z2{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline z2 operator-(z2 a, z2 b) { return z2{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline z2 operator*(z2 a, z2 b) {
  return z2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__host__ __device__ inline z2 operator*(double s, z2 a) { return z2{s * a.x, s * a.y}; }
__host__ __device__ inline z2 conjz(z2 a) { return z2{a.x, -a.y}; }
__host__ __device__ inline double abs2(z2 a) { return a.x * a.x + a.y * a.y; }
// a * conj(b)
__host__ __device__ inline z2 mulc(z2 a, z2 b) {
  return z2{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
// conj(a) * b
__host__ __device__ inline z2 cmul(z2 a, z2 b) {
  return z2{a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}
// element conversions between FP64 (z2) and FP32 (float2) complex storage
__device__ inline z2 widen(const z2 a) { return a; }
__device__ inline z2 widen(const float2 a) { return mk((double)a.x, (double)a.y); }
template <typename E> __device__ inline E narrow(z2 a);
template <> __device__ inline z2 narrow<z2>(z2 a) { return a; }
template <> __device__ inline float2 narrow<float2>(z2 a) {
  return make_float2((float)a.x, (float)a.y);
}

// ---------------------------------------------------------------- pair functions
struct PairParams {
  double a0, r0, d0, xi, tz, pi;
  int nscreen, nlayers;
};

// Slater-Koster pz hopping t(r_i, r_j), TightBinding.f90:216-230.
__host__ __device__ inline double pair_ft(const PairParams& p, double xi_, double yi, double zi,
                                          double xj, double yj, double zj) {
  double dx = xi_ - xj, dy = yi - yj, dz = zi - zj;
  double r = sqrt(dx * dx + dy * dy + dz * dz);
  if (!(r > 0.001)) return 0.0;
  double cs = dz / r;
  return -2.7 * exp((p.a0 - r) / p.r0) * (1.0 - cs * cs) + 0.48 * exp((p.d0 - r) / p.r0) * cs * cs;
}

// Gate-screened Coulomb v(r_i, r_j), HartreeFock.f90:1232-1274.
__host__ __device__ inline double pair_fv(const PairParams& p, double xi_, double yi, double zi,
                                          double xj, double yj, double zj) {
  double dx = xi_ - xj, dy = yi - yj, dz = zi - zj;
  double r = sqrt(dx * dx + dy * dy + dz * dz);
  double r2 = sqrt(dx * dx + dy * dy);
  double v = 0.0;
  if (p.nscreen == 2) {
    if (r > 0.01) {
      if (r2 < p.xi || r2 < 2.0 * (double)p.nlayers * p.tz) {
        double rho2 = (xj - xi_) * (xj - xi_) + (yj - yi) * (yj - yi);
        for (int n = -20; n <= 20; ++n) {
          double sgn = (n & 1) ? -1.0 : 1.0;
          double zz = zj - ((double)n * p.xi + zi * sgn);
          v += sgn / sqrt(rho2 + zz * zz);
        }
      } else {
        v = 2.0 * sqrt(2.0) / sqrt(r2 / p.xi) * exp(-p.pi * r2 / p.xi) / p.xi;
      }
    }
  } else if (p.nscreen == 1) {
    if (r > 0.01) {
      double zz = zi + zj + p.xi;
      v = 1.0 / r - 1.0 / sqrt(dx * dx + dy * dy + zz * zz);
    }
  }
  return v;
}

inline PairParams pair_params(const forge_params& P) {
  PairParams q;
  q.a0 = P.a0; q.r0 = P.r0; q.d0 = P.d0; q.xi = P.xi; q.tz = P.tz;
  q.pi = 3.14159265358979323846264338327950288;
  q.nscreen = (int)P.nscreen; q.nlayers = (int)P.nlayers;
  return q;
}

inline unsigned cdiv(i64 a, i64 b) { return (unsigned)((a + b - 1) / b); }

}  // namespace forge
