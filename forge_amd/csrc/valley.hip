// Valley projection (nenforceValley = 1): ValleyPhase and ValleyTransform,
// TightBinding.f90:60-211, for the doubled window density of
// Main.f90:469-494.
//
// Per k: the valley operator V(k) (n x n; nonzero only within one sublattice
// half of one layer) from ftvalley (TightBinding.f90:235-270), over the home
// cell and both images in every other stored cell; its projection
// P = C^H V C onto the window's nb eigenvectors (two zgemm); then, for all k
// at once, P = U Lambda U^H by the batched eigensolver of the Hamiltonian
// (reduction, divide and conquer, back-transform at order nb), and
// C <- C U sign(Lambda) U^H.  V is built one k at a time into an n x n
// scratch matrix.

#include "kernels.hpp"

namespace forge {

#define CK(x)               \
  do {                      \
    int _r = (x);           \
    if (_r != 0) return _r; \
  } while (0)

static const double VPI = 3.14159265358979323846264338327950288;

// i sign(cos 3 phi) / sqrt(27) for 0.2 < |d| < 1.2, phi the angle of R d
// (R row-major); the angle follows the reference's quadrant rules.
__device__ inline z2 ftvalley_d(double dx, double dy, const double* R) {
  const double rij = sqrt(dx * dx + dy * dy);
  if (!(rij < 1.2 && rij > 0.2)) return mk(0.0, 0.0);
  const double x = R[0] * dx + R[1] * dy, y = R[2] * dx + R[3] * dy;
  double phi;
  if (x > 0) phi = atan(y / x);
  else if (x < 0 && y > 0) phi = atan(y / x) + VPI;
  else phi = -VPI + atan(y / x);
  if (x == 0) phi = y > 0 ? VPI / 2.0 : -VPI / 2.0;
  return mk(0.0, copysign(1.0, cos(3.0 * phi)) / sqrt(27.0));
}

// V(i, j) for one k (fphase = 0), one thread per element, column j per block row.
__global__ void __launch_bounds__(256) k_valley_phase(ValleyGeom g, double vkx, double vky, z2* V) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  if (i >= g.n) return;
  z2 v = mk(0.0, 0.0);
  const i64 li = i / g.per, lj = j / g.per;
  const i64 hi = (i - li * g.per) < g.per / 2, hj = (j - lj * g.per) < g.per / 2;
  const int r = (li == lj && hi == hj && li < VALLEY_MAXL) ? g.rot[li] : 0;
  if (r == 1 || r == -1) {
    const double* R = r == 1 ? g.Rp : g.Rm;
    const double xi = g.X[i], yi = g.Y[i], xj = g.X[j], yj = g.Y[j];
    v = ftvalley_d(xi - xj, yi - yj, R);
    for (int c = 1; c < g.nnc; ++c) {
      const double Tx = g.cells.n1[c] * g.cells.t1x + g.cells.n2[c] * g.cells.t2x;
      const double Ty = g.cells.n1[c] * g.cells.t1y + g.cells.n2[c] * g.cells.t2y;
      const double ph = vkx * Tx + vky * Ty;
      const z2 e = mk(cos(ph), -sin(ph));  // e^{-i k.T}
      v = v + ftvalley_d(xi - (xj - Tx), yi - (yj - Ty), R) * e;
      v = v + ftvalley_d((xi - Tx) - xj, (yi - Ty) - yj, R) * conjz(e);
    }
  }
  V[i + g.n * j] = v;
}

int valley_phase(const ValleyGeom& g, double vkx, double vky, z2* V, hipStream_t s) {
  if (g.n > 65535) { set_error("valley_phase: n above the grid's y limit"); return -2; }
  hipLaunchKernelGGL(k_valley_phase, dim3((unsigned)cdiv(g.n, 256), (unsigned)g.n), dim3(256), 0, s,
                     g, vkx, vky, V);
  FORGE_LAUNCH();
  return 0;
}

// S_k(a, b) = sum_m U_k(a, m) sign(w_k(m)) conj(U_k(b, m)), nb x nb per k
__global__ void k_valley_sign(i64 nb, int K, const z2* __restrict__ U, const double* __restrict__ w,
                              z2* __restrict__ S) {
  const i64 q = (i64)blockIdx.x * 256 + threadIdx.x;
  if (q >= nb * nb * K) return;
  const i64 k = q / (nb * nb), r = q - k * nb * nb, a = r % nb, b = r / nb;
  const z2* Uk = U + k * nb * nb;
  const double* wk = w + k * nb;
  z2 acc = mk(0.0, 0.0);
  for (i64 m = 0; m < nb; ++m) acc = acc + copysign(1.0, wk[m]) * mulc(Uk[a + nb * m], Uk[b + nb * m]);
  S[q] = acc;
}

// workspace layout (bytes, 256-aligned pieces)
struct ValleyLayout {
  size_t D, Hs, trd, dc, Zq, ev, U, S, bt, T, total;
};
static ValleyLayout vlayout(i64 n, i64 nb, int K) {
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const i64 lds = (nb + 15) & ~(i64)15;
  ValleyLayout L{};
  size_t o = 0;
  L.D = o;   o += up(sizeof(z2) * n * nb);
  L.Hs = o;  o += up(sizeof(z2) * lds * nb * K);
  L.trd = o; o += up(tridiag_workspace_bytes(nb, K));
  L.dc = o;  o += up(dc_workspace_bytes(nb, K, nb));
  L.Zq = o;  o += up(sizeof(double) * nb * nb * K);
  L.ev = o;  o += up(sizeof(double) * nb * K);
  L.U = o;   o += up(sizeof(z2) * nb * nb * K);
  L.S = o;   o += up(sizeof(z2) * nb * nb * K);
  L.bt = o;  o += up(backtr_workspace_bytes(nb, K, nb));
  L.T = o;   o += up(sizeof(z2) * n * nb * K);
  L.total = o;
  return L;
}

size_t valley_workspace_bytes(i64 n, i64 nb, int K) { return vlayout(n, nb, K).total; }

// C (n x nb per k, ld n, stride n nb) <- C sign(C^H V_k C) for the K momenta
// vk (host, [K][2]).  V: n x n scratch.
int valley_transform(const ValleyGeom& g, const double* vk, int K, i64 nb, z2* C, z2* V,
                     void* work, hipStream_t s) {
  const i64 n = g.n;
  const ValleyLayout L = vlayout(n, nb, K);
  char* b = (char*)work;
  z2* D = (z2*)(b + L.D);
  z2* Hs = (z2*)(b + L.Hs);
  const i64 lds = (nb + 15) & ~(i64)15;
  for (int k = 0; k < K; ++k) {
    CK(valley_phase(g, vk[2 * k], vk[2 * k + 1], V, s));
    const z2* Ck = C + (size_t)k * n * nb;
    // D = V C_k, then P_k = C_k^H D (lower triangle read by the eigensolver)
    CK(zgemm(Op::N, Op::N, n, nb, n, 1.0, V, n, 0, Ck, n, 0, 0.0, D, n, 0, 1, s));
    CK(zgemm(Op::H, Op::N, nb, nb, n, 1.0, Ck, n, 0, D, n, 0, 0.0, Hs + (size_t)k * lds * nb, lds, 0,
             1, s));
  }
  // P_k = U_k Lambda_k U_k^H for all k
  TridiagWork t{};
  CK(tridiag_bind(t, b + L.trd, nb, K));
  t.A = Hs;
  t.lda = lds;
  t.sA = (size_t)lds * nb;
  CK(tridiag(t, s, nullptr, nullptr, nullptr));
  double* Zq = (double*)(b + L.Zq);
  double* ev = (double*)(b + L.ev);
  const int fails = dc_solve(t.d, t.e, nb, K, nb, ev, Zq, b + L.dc, s);
  if (fails != 0) {
    set_error("valley_transform: the divide-and-conquer solve of a projected valley operator failed");
    return fails < 0 ? fails : -1;
  }
  z2* U = (z2*)(b + L.U);
  CK(backtransform(t, Zq, nb, (size_t)nb * nb, nb, U, nb, (size_t)nb * nb, b + L.bt, s));
  z2* S = (z2*)(b + L.S);
  hipLaunchKernelGGL(k_valley_sign, dim3((unsigned)cdiv(nb * nb * K, 256)), dim3(256), 0, s, nb, K, U,
                     ev, S);
  FORGE_LAUNCH();
  // C_k <- C_k S_k
  z2* T = (z2*)(b + L.T);
  CK(zgemm(Op::N, Op::N, n, nb, nb, 1.0, C, n, (size_t)n * nb, S, nb, (size_t)nb * nb, 0.0, T, n,
           (size_t)n * nb, K, s));
  FORGE_HIP(hipMemcpyAsync(C, T, sizeof(z2) * n * nb * K, hipMemcpyDeviceToDevice, s));
  return 0;
}

#undef CK

}  // namespace forge
