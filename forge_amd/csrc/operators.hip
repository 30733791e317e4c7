// Pair tables, Hartree kernel, batched H(k) assembly and the energy pass.
//
//   k_pair_tables   V_c(i,j) = fv(r_i, r_j - T_c), T_c(i,j) = ft(r_i, r_j - T_c) for
//                   every stored cell c (c = 0 is the home cell); fv/ft follow
//                   HartreeFock.f90:1232-1274 / TightBinding.f90:216-230.  Both are
//                   symmetric and in-plane translation invariant, so
//                   fv(r_i, r_j + T_c) = V_c(j,i) and ft(r_i, r_j + T_c) = T_c(j,i).
//   k_longrange     LongRangeInteraction, HartreeFock.f90:14-53.
//   k_assemble      HamiltonianHartreeFock, HartreeFock.f90:57-127, for a batch of
//                   k-points in one pass over rho and the tables:
//                     H_k(i,j) = W_0(i,j) + sum_c [A_c(i,j) e^{-i k.T_c} + conj(A_c(j,i)) e^{+i k.T_c}]
//                     W_0 = T_0 - alpha V_0 conj(rho_0),  A_c = T_c - alpha V_c conj(rho_c)
//                   plus Potential and the layer bias on the diagonal (lower triangle).
//   k_energy_pass   GetKineticEnergy / GetFockEnergy (HartreeFock.f90:1012-1145)
//                   for two densities a (rho_in) and b (rho_new) and their cross term
//                   in one read of rho_in, rho_new and the tables.
#include "common.hpp"
#include "kernels.hpp"

namespace forge {

// ------------------------------------------------------------------ tables
__global__ void __launch_bounds__(256) k_pair_tables(i64 n, int nnc, const double* __restrict__ X,
                                                     const double* __restrict__ Y,
                                                     const double* __restrict__ Z, CellShifts cs,
                                                     PairParams pp, double* __restrict__ V,
                                                     double* __restrict__ T) {
  i64 i = (i64)blockIdx.x * 16 + (threadIdx.x & 15);
  i64 j = (i64)blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i >= n || j >= n) return;
  double xi = X[i], yi = Y[i], zi = Z[i], xj = X[j], yj = Y[j], zj = Z[j];
  size_t nn = (size_t)n * n;
  for (int c = 0; c < nnc; ++c) {
    // r_j - n1 t1 - n2 t2 evaluated left to right as in HartreeFock.f90:91
    double sx = (xj - cs.n1[c] * cs.t1x) - cs.n2[c] * cs.t2x;
    double sy = (yj - cs.n1[c] * cs.t1y) - cs.n2[c] * cs.t2y;
    V[c * nn + i + (size_t)n * j] = pair_fv(pp, xi, yi, zi, sx, sy, zj);
    T[c * nn + i + (size_t)n * j] = pair_ft(pp, xi, yi, zi, sx, sy, zj);
  }
}

int pair_tables(const DeviceGeom& g, const PairParams& pp, double* V, double* T,
                hipStream_t s) {
  dim3 grid(cdiv(g.n, 16), cdiv(g.n, 16));
  hipLaunchKernelGGL(k_pair_tables, grid, dim3(256), 0, s, g.n, g.nnc, g.X, g.Y, g.Z, g.cells,
                     pp, V, T);
  FORGE_LAUNCH();
  return 0;
}

// ------------------------------------------------------------------ LongRange
// One thread per (a, b) with b >= a; coordinates of a 64-wide column tile are
// staged in LDS.
__global__ void __launch_bounds__(256) k_longrange(i64 n, const double* __restrict__ X,
                                                   const double* __restrict__ Y,
                                                   const double* __restrict__ Z, double t1x,
                                                   double t1y, double t2x, double t2y, int numC,
                                                   PairParams pp, double* __restrict__ L) {
  __shared__ double sx[16], sy[16], sz[16];
  i64 a = (i64)blockIdx.y * 16 + (threadIdx.x >> 4);
  i64 b = (i64)blockIdx.x * 16 + (threadIdx.x & 15);
  if (threadIdx.x < 16) {
    i64 bb = (i64)blockIdx.x * 16 + threadIdx.x;
    sx[threadIdx.x] = bb < n ? X[bb] : 0.0;
    sy[threadIdx.x] = bb < n ? Y[bb] : 0.0;
    sz[threadIdx.x] = bb < n ? Z[bb] : 0.0;
  }
  __syncthreads();
  if (a >= n || b >= n || b < a) return;
  double xa = X[a], ya = Y[a], za = Z[a];
  double xb = sx[threadIdx.x & 15], yb = sy[threadIdx.x & 15], zb = sz[threadIdx.x & 15];
  double s = 0.0;
  for (int n1 = -numC; n1 <= numC; ++n1)
    for (int n2 = -numC; n2 <= numC; ++n2)
      if (n1 + n2 <= numC && n1 + n2 >= -numC) {
        double px = (xa - n1 * t1x) - n2 * t2x;
        double py = (ya - n1 * t1y) - n2 * t2y;
        s += pair_fv(pp, px, py, za, xb, yb, zb);
      }
  L[a + (size_t)n * b] = s;
  L[b + (size_t)n * a] = s;
}

int longrange(const DeviceGeom& g, const PairParams& pp, int numC, double* L, hipStream_t s) {
  dim3 grid(cdiv(g.n, 16), cdiv(g.n, 16));
  hipLaunchKernelGGL(k_longrange, grid, dim3(256), 0, s, g.n, g.X, g.Y, g.Z, g.cells.t1x,
                     g.cells.t1y, g.cells.t2x, g.cells.t2y, numC, pp, L);
  FORGE_LAUNCH();
  return 0;
}

// ------------------------------------------------------------------ assembly
constexpr int AT = 32;        // assembly tile edge
constexpr int ASM_CELLS = 4;  // stored cells of numI = 1 (Geometry.f90:14-35)

// grid: (lower-triangle tile index); block 256 threads: thread (tx, ty) with
// tx in [0,32) rows, ty in [0,8) column groups of 4.
__global__ void __launch_bounds__(256) k_assemble(i64 n, int nnc, int K,
                                                  const z2* __restrict__ rho,
                                                  const double* __restrict__ V,
                                                  const double* __restrict__ T,
                                                  const double* __restrict__ pot,
                                                  const double* __restrict__ bias, double alpha,
                                                  const double* __restrict__ phase,  // [K][nnc][2]
                                                  z2* __restrict__ H, size_t ldH,
                                                  size_t strideH) {
  __shared__ z2 tr[ASM_CELLS - 1][AT][AT + 1];  // image cells 1..nnc-1 (<= 50 KB)
  // decode lower-triangle tile (I >= J) from the linear block index
  i64 t = blockIdx.x;
  i64 I = (i64)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= t) ++I;
  while (I * (I + 1) / 2 > t) --I;
  i64 J = t - I * (I + 1) / 2;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const size_t nn = (size_t)n * n;
  const i64 i0 = I * AT, j0 = J * AT;
  // transposed partner tiles: element (j, i) for j in tile J, i in tile I
  for (int c = 1; c < nnc; ++c) {
    for (int q = 0; q < 4; ++q) {
      int cc = ty * 4 + q;  // column within tile I
      i64 jr = j0 + tx, ic = i0 + cc;
      z2 a = mk(0.0, 0.0);
      if (jr < n && ic < n) {
        size_t o = jr + (size_t)n * ic;
        double tv = T[c * nn + o], vv = V[c * nn + o];
        z2 r = rho ? rho[c * nn + o] : mk(0.0, 0.0);
        a = mk(tv - alpha * vv * r.x, alpha * vv * r.y);  // T - alpha V conj(rho)
      }
      tr[c - 1][tx][cc] = a;  // tr[c-1][j - j0][i - i0]
    }
  }
  __syncthreads();
  for (int q = 0; q < 4; ++q) {
    int cc = ty * 4 + q;
    i64 i = i0 + tx, j = j0 + cc;
    if (i >= n || j >= n || j > i) continue;
    size_t o = i + (size_t)n * j;
    z2 r0 = rho ? rho[o] : mk(0.0, 0.0);
    double v0 = V[o], t0 = T[o];
    z2 w0 = mk(t0 - alpha * v0 * r0.x, alpha * v0 * r0.y);
    if (i == j) w0.x += pot[i] + bias[i];
    // the image cells in registers: fixed-size arrays under fully unrolled
    // loops (a runtime bound would index them dynamically, in scratch memory)
    z2 a[ASM_CELLS], b[ASM_CELLS];
#pragma unroll
    for (int c = 1; c < ASM_CELLS; ++c) {
      a[c] = mk(0.0, 0.0);
      b[c] = mk(0.0, 0.0);
      if (c < nnc) {
        double tv = T[c * nn + o], vv = V[c * nn + o];
        z2 r = rho ? rho[c * nn + o] : mk(0.0, 0.0);
        a[c] = mk(tv - alpha * vv * r.x, alpha * vv * r.y);
        b[c] = conjz(tr[c - 1][cc][tx]);  // conj(A_c(j, i))
      }
    }
    for (int k = 0; k < K; ++k) {
      z2 h = w0;
#pragma unroll
      for (int c = 1; c < ASM_CELLS; ++c) {
        if (c >= nnc) continue;
        double co = phase[(k * nnc + c) * 2], si = phase[(k * nnc + c) * 2 + 1];
        // a e^{-i th} + b e^{+i th}
        h.x += a[c].x * co + a[c].y * si + b[c].x * co - b[c].y * si;
        h.y += a[c].y * co - a[c].x * si + b[c].y * co + b[c].x * si;
      }
      H[k * strideH + i + ldH * j] = h;
    }
  }
}

int assemble(i64 n, int nnc, int K, const z2* rho, const double* V, const double* T,
             const double* pot, const double* bias, double alpha, const double* phase, z2* H,
             size_t ldH, size_t strideH, hipStream_t s) {
  if (nnc > ASM_CELLS) { set_error("assemble: at most 4 stored cells (numI = 1)"); return -2; }
  i64 nt = (n + AT - 1) / AT;
  i64 tiles = nt * (nt + 1) / 2;
  hipLaunchKernelGGL(k_assemble, dim3((unsigned)tiles), dim3(256), 0, s, n, nnc, K, rho, V, T,
                     pot, bias, alpha, phase, H, ldH, strideH);
  FORGE_LAUNCH();
  return 0;
}

// ------------------------------------------------------------------ energies
// Per-block partial sums (fixed order, reduced by k_reduce_rows):
//   [0] K_a  [1] K_b  [2] F_aa  [3] F_ab  [4] F_bb
// K = sum T_0 Re rho_0 + 2 sum_c T_c Re rho_c
// F(x,y) = -1/2 sum [V_0 Re(conj x_0 y_0) + 2 sum_c V_c Re(conj x_c y_c)]
__global__ void __launch_bounds__(256) k_energy_pass(size_t nn, int nnc,
                                                     const z2* __restrict__ a,
                                                     const z2* __restrict__ b,
                                                     const double* __restrict__ V,
                                                     const double* __restrict__ T,
                                                     double* __restrict__ part) {
  double acc[5] = {0, 0, 0, 0, 0};
  for (size_t o = (size_t)blockIdx.x * 256 + threadIdx.x; o < nn; o += (size_t)gridDim.x * 256) {
    for (int c = 0; c < nnc; ++c) {
      double w = c ? 2.0 : 1.0;
      double tv = T[c * nn + o], vv = V[c * nn + o];
      z2 x = a[c * nn + o];
      acc[0] += w * tv * x.x;
      acc[2] += -0.5 * w * vv * abs2(x);
      if (b) {
        z2 y = b[c * nn + o];
        acc[1] += w * tv * y.x;
        acc[3] += -0.5 * w * vv * (x.x * y.x + x.y * y.y);
        acc[4] += -0.5 * w * vv * abs2(y);
      }
    }
  }
  __shared__ double red[5][256];
  for (int q = 0; q < 5; ++q) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int q = 0; q < 5; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 5; ++q) part[(size_t)blockIdx.x * 5 + q] = red[q][0];
}

// Hartree matvec for two density vectors in one read of L:
//   ya = L da, yb = L db (rows a of L; L symmetric, column-major).
__global__ void __launch_bounds__(256) k_hartree_matvec(i64 n, const double* __restrict__ L,
                                                        const double* __restrict__ da,
                                                        const double* __restrict__ db,
                                                        double* __restrict__ ya,
                                                        double* __restrict__ yb) {
  i64 r = (i64)blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  double sa = 0, sb = 0;
  for (i64 m = 0; m < n; ++m) {
    double l = L[r + (size_t)n * m];
    sa += l * da[m];
    if (db) sb += l * db[m];
  }
  ya[r] = sa;
  if (db) yb[r] = sb;
}

// Deterministic reduction of P partial rows of width W: out[q] = sum_p part[p*W+q].
__global__ void k_reduce_rows(const double* __restrict__ part, int P, int W,
                              double* __restrict__ out) {
  int q = threadIdx.x;
  if (q >= W) return;
  double s = 0;
  for (int p = 0; p < P; ++p) s += part[(size_t)p * W + q];
  out[q] = s;
}

int energy_pass(size_t nn, int nnc, const z2* a, const z2* b, const double* V, const double* T,
                double* part, int nblocks, double* out5, hipStream_t s) {
  hipLaunchKernelGGL(k_energy_pass, dim3(nblocks), dim3(256), 0, s, nn, nnc, a, b, V, T, part);
  FORGE_LAUNCH();
  hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(64), 0, s, part, nblocks, 5, out5);
  FORGE_LAUNCH();
  return 0;
}

int hartree_matvec(i64 n, const double* L, const double* da, const double* db, double* ya,
                   double* yb, hipStream_t s) {
  hipLaunchKernelGGL(k_hartree_matvec, dim3(cdiv(n, 256)), dim3(256), 0, s, n, L, da, db, ya, yb);
  FORGE_LAUNCH();
  return 0;
}

// ------------------------------------------------------------------ mixing
// rho_in <- (1 - s) rho_in + s rho_new over all stored cells (Main.f90:558).
__global__ void k_mix(size_t tot, z2* __restrict__ a, const z2* __restrict__ b, double s) {
  for (size_t o = (size_t)blockIdx.x * 256 + threadIdx.x; o < tot; o += (size_t)gridDim.x * 256) {
    z2 x = a[o], y = b[o];
    a[o] = mk(s * y.x + (1.0 - s) * x.x, s * y.y + (1.0 - s) * x.y);
  }
}

int mix(size_t tot, z2* a, const z2* b, double s, hipStream_t st) {
  hipLaunchKernelGGL(k_mix, dim3(2048), dim3(256), 0, st, tot, a, b, s);
  FORGE_LAUNCH();
  return 0;
}

// dens[i] = Re rho_0(i,i) - 1/2
__global__ void k_site_density(i64 n, const z2* __restrict__ rho, double* __restrict__ d) {
  i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) d[i] = rho[i + (size_t)n * i].x - 0.5;
}

int site_density(i64 n, const z2* rho, double* d, hipStream_t s) {
  hipLaunchKernelGGL(k_site_density, dim3(cdiv(n, 256)), dim3(256), 0, s, n, rho, d);
  FORGE_LAUNCH();
  return 0;
}

}  // namespace forge
