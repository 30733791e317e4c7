// C3-reduced k-set (nenforceC3 = 1): Solve_C3 and GetFock_C3,
// HartreeFock.f90:131-398 and 564-881, with C3_RelatedPoints,
// Geometry.f90:1172-1223.
//
// C3 takes momentum k = (i1, i2) to k2 = (-i1-i2, i1) and k3 = (i2, -i1-i2)
// (mod numk).  Only the first momentum of each orbit is diagonalised.  Every
// other object the reference derives from it is a permuted copy with extra
// phases on the two cell-corner sites of each layer:
//   image column  w(i)   = v(p^r(i)) f(i)                    r = 1, 2
//   image block   T'(i,j) = T(p^r(i), p^r(j)) g(i) conj(g(j))
// p is the site permutation of the +120 degree rotation within each layer and
// sublattice half, and f, g are 1 off the corners.  Both kernels are gathers
// bound by HBM, with one pass over their output.

#include <vector>

#include "kernels.hpp"

namespace forge {

// Corner factor of site i in factor set `set` (-1: none).  tab holds, per
// set and layer, the factors of the first corner (last site of the first
// sublattice half) and the second corner (last site of the layer).
__device__ inline z2 c3_corner(const C3Geom& g, int set, i64 i) {
  if (set < 0) return mk(1.0, 0.0);
  const i64 l = i / g.per, r = i - l * g.per;
  if (r == g.per / 2 - 1) return g.tab[((size_t)set * g.nl + l) * 2];
  if (r == g.per - 1) return g.tab[((size_t)set * g.nl + l) * 2 + 1];
  return mk(1.0, 0.0);
}

// dst(:, d.dst) = src(p^d.r(i), d.src) * corner(d.set, i), one column per
// descriptor.  src and dst may be one buffer when no destination column is
// also a source column (checked by the host).
__global__ void __launch_bounds__(256) k_c3_columns(C3Geom g, i64 ncol, const C3Col* __restrict__ d,
                                                    const z2* src, z2* dst) {
  const i64 q = (i64)blockIdx.x * 256 + threadIdx.x;
  if (q >= g.n * ncol) return;
  const i64 i = q % g.n, c = q / g.n;
  const C3Col e = d[c];
  const i64 si = e.r == 0 ? i : (i64)g.perm[(size_t)(e.r - 1) * g.n + i];
  dst[i + g.n * e.dst] = src[si + g.n * e.src] * c3_corner(g, e.set, i);
}

int c3_columns(const C3Geom& g, i64 ncol, const C3Col* d, const z2* src, z2* dst, hipStream_t s) {
  if (ncol <= 0) return 0;
  hipLaunchKernelGGL(k_c3_columns, dim3((unsigned)cdiv(g.n * ncol, 256)), dim3(256), 0, s, g, ncol, d,
                     src, dst);
  FORGE_LAUNCH();
  return 0;
}

__device__ inline z2 herm_lower(const z2* R, i64 n, i64 a, i64 b) {
  return a >= b ? R[a + n * b] : conjz(R[b + n * a]);
}

// rho_c(i,j) (+)= w [ph0_c T(i,j) + ph1_c T(p i, p j) f(i) conj(f(j))
//                    + ph2_c T(pp i, pp j) u(i) conj(u(j))],   T = conj(R),
// R Hermitian with its lower triangle valid.  The three terms are the orbit's
// bulk block at k and its C3 and C3^2 images (HartreeFock.f90:284-389).
__global__ void __launch_bounds__(256) k_c3_bulk(C3Geom g, int nnc, const z2* __restrict__ R,
                                                 const double* __restrict__ ph0,
                                                 const double* __restrict__ ph1,
                                                 const double* __restrict__ ph2, int fset, int uset,
                                                 double w, z2* __restrict__ rho, int first) {
  const i64 n = g.n;
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  if (i >= n) return;
  const int* p1 = g.perm;
  const int* p2 = g.perm + n;
  const z2 t0 = conjz(herm_lower(R, n, i, j));
  const z2 t1 = conjz(herm_lower(R, n, p1[i], p1[j])) *
                mulc(c3_corner(g, fset, i), c3_corner(g, fset, j));
  const z2 t2 = conjz(herm_lower(R, n, p2[i], p2[j])) *
                mulc(c3_corner(g, uset, i), c3_corner(g, uset, j));
  const size_t nn = (size_t)n * n, o = i + n * j;
  for (int c = 0; c < nnc; ++c) {
    const z2 add = w * (mk(ph0[2 * c], ph0[2 * c + 1]) * t0 + mk(ph1[2 * c], ph1[2 * c + 1]) * t1 +
                        mk(ph2[2 * c], ph2[2 * c + 1]) * t2);
    rho[c * nn + o] = first ? add : rho[c * nn + o] + add;
  }
}

// Bulk density of K orbit representatives solved in one batch (columns of C,
// ld ldC, stride sC; bands 1..nocc): per representative one rank-nocc HERK
// into the scratch matrix after the nnc cells of rho, then its three images
// into every cell.  accumulate = false overwrites at the first representative.
int bulk_density_c3(const C3Geom& g, int K, const z2* C, size_t ldC, size_t sC, i64 nocc, int nnc,
                    const C3Slot* slot, z2* rho, bool accumulate, hipStream_t s) {
  const i64 n = g.n;
  const size_t nn = (size_t)n * n;
  z2* R = rho + nnc * nn;
  if (n > 65535) { set_error("bulk_density_c3: n above the grid's y limit"); return -2; }
  dim3 grid((unsigned)cdiv(n, 256), (unsigned)n);
  for (int k = 0; k < K; ++k) {
    if (nocc > 0) {
      const int r = herk_lower(n, nocc, C + k * sC, (i64)ldC, R, n, s);
      if (r) return r;
    } else {
      FORGE_HIP(hipMemsetAsync(R, 0, sizeof(z2) * nn, s));
    }
    const C3Slot& e = slot[k];
    hipLaunchKernelGGL(k_c3_bulk, grid, dim3(256), 0, s, g, nnc, R, e.ph0, e.ph1, e.ph2, e.fset,
                       e.uset, e.w, rho, (k == 0 && !accumulate) ? 1 : 0);
    FORGE_LAUNCH();
  }
  return 0;
}

// C3_RelatedPoints, Geometry.f90:1172-1223, on the host (set-up, once per
// geometry): perm[i] (0-based) = the site of the same layer and sublattice
// half within 1e-3 of C3 r_i; the two corners of each layer map to
// themselves.  Unmatched sites stay -1.  Returns the number of sites assigned.
i64 c3_pairs(const double* C, i64 n, i64 nl, i64* perm) {
  const double h3 = 0.5 * std::sqrt(3.0);
  const i64 per = n / nl, half = per / 2;
  i64 cnt = 0;
  for (i64 i = 0; i < n; ++i) perm[i] = -1;
  for (i64 l = 0; l < nl; ++l)
    for (int hf = 0; hf < 2; ++hf) {
      const i64 lo = l * per + hf * half, hi = lo + half;
      for (i64 i = lo; i < hi; ++i) {
        const double x = C[i], y = C[i + n];
        const double x3 = -0.5 * x - h3 * y, y3 = h3 * x - 0.5 * y;
        for (i64 j = lo; j < hi; ++j) {
          const double dx = C[j] - x3, dy = C[j + n] - y3;
          if (std::sqrt(dx * dx + dy * dy) < 1e-3) { perm[i] = j; ++cnt; }
        }
      }
    }
  for (i64 l = 0; l < nl; ++l) {
    perm[l * per + half - 1] = l * per + half - 1;
    perm[l * per + per - 1] = l * per + per - 1;
    cnt += 2;
  }
  return cnt;
}

}  // namespace forge
