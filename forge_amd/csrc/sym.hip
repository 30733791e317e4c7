// Symmetry enforcement on each iteration's new density (nenforceT /
// nenforceC2 / nenforceC2T, Main.f90:514-534) and C2_RelatedPoints
// (Geometry.f90:1140-1167).
//
// In the reference's order, on every stored cell:
//   T    rho <- Re rho
//   C2   rho(i,j) <- (rho(i,j) + conj(rho(p(j), p(i)))) / 2
//   C2T  rho(i,j) <- (rho(i,j) + rho(p(j), p(i))) / 2
// with p the C2 partner of each site.  C2 and C2T read a scratch copy of the
// cell, since each output reads the mirrored element.  One pass over n^2
// elements per cell: a gather bound by HBM.
#include "kernels.hpp"

namespace forge {

__global__ void __launch_bounds__(256) k_sym_real(size_t count, z2* rho) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < count) rho[q].y = 0.0;
}

__global__ void __launch_bounds__(256) k_sym_c2(i64 n, const z2* __restrict__ src, z2* __restrict__ dst,
                                                const int* __restrict__ p, int conj) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  if (i >= n) return;
  z2 o = src[(i64)p[j] + n * (i64)p[i]];
  if (conj) o = conjz(o);
  dst[i + n * j] = 0.5 * (src[i + n * j] + o);
}

int symmetrize(i64 n, int nnc, z2* rho, z2* scratch, int sym, const int* perm, hipStream_t s) {
  const size_t nn = (size_t)n * n;
  if (sym & 1) {
    const size_t count = nn * nnc;
    hipLaunchKernelGGL(k_sym_real, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, count, rho);
    FORGE_LAUNCH();
  }
  if (!(sym & 6)) return 0;
  if (!perm) { set_error("symmetrize: C2 or C2T without C2 partners"); return -2; }
  if (n > 65535) { set_error("symmetrize: n above the grid's y limit"); return -2; }
  const dim3 grid((unsigned)cdiv(n, 256), (unsigned)n);
  for (int f = 2; f <= 4; f *= 2) {
    if (!(sym & f)) continue;
    for (int c = 0; c < nnc; ++c) {
      FORGE_HIP(hipMemcpyAsync(scratch, rho + c * nn, sizeof(z2) * nn, hipMemcpyDeviceToDevice, s));
      hipLaunchKernelGGL(k_sym_c2, grid, dim3(256), 0, s, n, scratch, rho + c * nn, perm, f == 2 ? 1 : 0);
      FORGE_LAUNCH();
    }
  }
  return 0;
}

// C2_RelatedPoints on the host (set-up): within each layer, site i of the
// first sublattice half and site j of the second are partners when
// r_i = -r_j in the plane (within 1e-5); perm (0-based) gets both directions,
// -1 where unmatched.  Returns the number of pairs.
i64 c2_pairs(const double* C, i64 n, i64 nl, i64* perm) {
  const i64 per = n / nl, half = per / 2;
  i64 cnt = 0;
  for (i64 i = 0; i < n; ++i) perm[i] = -1;
  for (i64 l = 0; l < nl; ++l)
    for (i64 i = l * per; i < l * per + half; ++i)
      for (i64 j = l * per + half; j < (l + 1) * per; ++j) {
        const double x = C[i] + C[j], y = C[i + n] + C[j + n];
        if (std::sqrt(x * x + y * y) < 1e-5) { perm[i] = j; perm[j] = i; ++cnt; }
      }
  return cnt;
}

}  // namespace forge
