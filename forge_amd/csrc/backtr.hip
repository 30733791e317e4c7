// Back-transformation C = Q Z of the tridiagonal eigenvectors Z (real, n x m)
// to eigenvectors of the Hermitian matrices, Q = H_0 ... H_{n-2} from
// tridiag.hip (zunmtr('L','L','N') semantics; zheevr applies it in
// LapackRoutines.f90:414-426).  Reflectors are applied in blocks of BT in the
// compact WY form  H_i0 ... H_{i0+b-1} = I - V T V^H  (T upper triangular,
// LAPACK zlarft 'F','C'), last block first:
//     Y = V^H C,  Y <- T Y,  C <- C - V Y    (matrix-core GEMMs of gemm.hip, batched over k)
// The clean V (unit diagonal, zeros above) and T are built by small kernels.
//
// FP32 variant (the mixed-precision phase of the SCF): the same blocks in
// complex FP32 (the FP32 matrix cores, V, T and C rounded to FP32),
// then every column is scaled to unit 2-norm in FP64 and written to the FP64
// output.  Eigenvector errors then enter the band energies only at second
// order: for psi = psi* + d, <psi|H|psi> / <psi|psi> - eps = O(|d|^2), while
// an unnormalised column would carry the first-order term 2 eps Re<psi*|d>.
//
// Block width autotuning: the width trades passes over C (n / BT reads and
// writes of the n x m eigenvectors) against the triangular work of G and T
// (growing with BT) and the efficiency of the matrix-core GEMMs at inner
// dimension BT.  On the first call for a shape (n, m, K, FP64/FP32) every
// candidate width applies the first block (the largest trailing order) to
// the promoted eigenvectors twice; the second run is timed with HIP events
// and the width with the least time per reflector is kept for the shape for
// the rest of the process.
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include "kernels.hpp"

namespace forge {

constexpr int BT_MAX = 256;
constexpr int BT_CAND[3] = {64, 128, 256};

size_t backtr_workspace_bytes(i64 n, int K, i64 m) {
  // V (n x BT), G (BT x BT), T (BT x BT), Y (BT x m), Y2 (BT x m) per matrix
  size_t per = (size_t)n * BT_MAX + 2 * BT_MAX * BT_MAX + 2 * (size_t)BT_MAX * m;
  return sizeof(z2) * per * K + 4096;
}

template <typename E>
__global__ void k_bt_promote(size_t nm, const double* __restrict__ Z, size_t ldZ, size_t sZ,
                             i64 n, i64 m, E* __restrict__ C, size_t ldC, size_t sC) {
  int k = blockIdx.y;
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < nm; q += (size_t)gridDim.x * 256) {
    i64 r = (i64)(q % n), c = (i64)(q / n);
    C[k * sC + r + ldC * c] = narrow<E>(mk(Z[k * sZ + r + ldZ * c], 0.0));
  }
}

// V(r', q) for rows r = i0+voff+r' of reflector j = i0+q: 1 at r = j+voff, A(r, j)
// below (voff = 1 after the one-stage reduction, the band width after the
// first stage of the two-stage one).
template <typename E>
__global__ void k_bt_vclean(i64 n, i64 i0, int nb, int voff, const z2* __restrict__ A, size_t sA,
                            i64 lda, E* __restrict__ V, size_t sV) {
  int k = blockIdx.y;
  i64 mt = n - i0 - voff;
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < (size_t)mt * nb;
       q += (size_t)gridDim.x * 256) {
    i64 rr = (i64)(q % mt), c = (i64)(q / mt);
    i64 r = i0 + voff + rr, j = i0 + c;
    z2 v;
    if (r == j + voff) v = mk(1.0, 0.0);
    else if (r > j + voff) v = A[k * sA + r + lda * j];
    else v = mk(0.0, 0.0);
    V[k * sV + rr + (size_t)n * c] = narrow<E>(v);
  }
}

// T (nb x nb, upper) from G = V^H V and tau (zlarft 'F','C').  T^-1 = striu(G)
// + diag(1/tau), so column i of T follows from the rows below it:
//   T(i,i) = tau_i,   T(r,i) = -tau_r sum_{c=r+1..i} G(r,c) T(c,i)   (r < i),
// which also holds for tau = 0 (an identity reflector: its row and column are
// zero).  One wave per column (grid nb x K), the column in LDS, G(r, c) read
// as conj(G(c, r)) (contiguous in c) one row ahead of its use.
template <typename E, int BT>
__global__ void __launch_bounds__(64) k_bt_tbuild(int nb, i64 i0, i64 n, const E* __restrict__ G,
                                                  const z2* __restrict__ tau, E* __restrict__ T) {
  constexpr int NL = BT / 64;  // entries of a row per lane
  const int i = blockIdx.x, k = blockIdx.y, l = threadIdx.x;
  const E* g = G + (size_t)k * BT * BT;
  E* t = T + (size_t)k * BT * BT;
  const z2* tk = tau + (size_t)k * n + i0;
  __shared__ z2 x[BT];
  for (int r = i + 1 + l; r < nb; r += 64) t[r + BT * i] = narrow<E>(mk(0.0, 0.0));
  if (l == 0) x[i] = tk[i];
  __syncthreads();
  // row r needs G(r, c) for c = r+1+l, r+65+l, ... (c <= i)
  auto load = [&](int r, int c) { return (r >= 0 && c <= i) ? conjz(widen(g[c + BT * r])) : mk(0.0, 0.0); };
  z2 gc[NL];
#pragma unroll
  for (int q = 0; q < NL; ++q) gc[q] = load(i - 1, i + 64 * q + l);
  for (int r = i - 1; r >= 0; --r) {
    z2 gn[NL];
#pragma unroll
    for (int q = 0; q < NL; ++q) gn[q] = load(r - 1, r + 64 * q + l);
    z2 p = mk(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < NL; ++q) {
      const int c = r + 1 + 64 * q + l;
      if (c <= i) p = p + gc[q] * x[c];
    }
    double sx = p.x, sy = p.y;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      sx += __shfl_xor(sx, off, 64);
      sy += __shfl_xor(sy, off, 64);
    }
    if (l == 0) x[r] = (-1.0) * (tk[r] * mk(sx, sy));
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NL; ++q) gc[q] = gn[q];
  }
  for (int r = l; r <= i; r += 64) t[r + BT * i] = narrow<E>(x[r]);
}

// C(:, j) = C32(:, j) / |C32(:, j)|_2 in FP64.  grid (m, K), 256 threads.
__global__ void __launch_bounds__(256) k_bt_normalise(i64 n, const float2* __restrict__ C32, size_t ldC32,
                                                      size_t sC32, z2* __restrict__ C, size_t ldC,
                                                      size_t sC) {
  const int k = blockIdx.y;
  const i64 j = blockIdx.x;
  const float2* c32 = C32 + k * sC32 + ldC32 * j;
  z2* c64 = C + k * sC + ldC * j;
  double ss = 0.0;
  for (i64 r = threadIdx.x; r < n; r += 256) {
    const float2 v = c32[r];
    ss += (double)v.x * v.x + (double)v.y * v.y;
  }
  __shared__ double red[256];
  red[threadIdx.x] = ss;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const double inv = red[0] > 0.0 ? 1.0 / sqrt(red[0]) : 0.0;
  for (i64 r = threadIdx.x; r < n; r += 256) c64[r] = inv * widen(c32[r]);
}

#define CK_BT(x)              \
  do {                        \
    int _r = (x);             \
    if (_r != 0) return _r;   \
  } while (0)

// the matrix-core GEMM of each storage type (gemm.hip)
static int gemm_sb(Op ta, Op tb, i64 M, i64 N, i64 Kd, double alpha, const z2* A, i64 lda,
                   size_t sA, const z2* B, i64 ldb, size_t sB, double beta, z2* C, i64 ldc,
                   size_t sC, int batch, hipStream_t s) {
  return zgemm(ta, tb, M, N, Kd, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, s);
}
static int gemm_sb(Op ta, Op tb, i64 M, i64 N, i64 Kd, double alpha, const float2* A, i64 lda,
                   size_t sA, const float2* B, i64 ldb, size_t sB, double beta, float2* C,
                   i64 ldc, size_t sC, int batch, hipStream_t s) {
  return cgemm(ta, tb, M, N, Kd, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, s);
}

// C (ldC, stride sC per matrix) <- Q Z, blocks of BT reflectors last to first,
// in storage type E.  first_only: C <- (first block) Z, for timing, with the
// event `promoted` recorded between the promotion of Z and the block.
template <typename E, int BT>
static int apply_q(const TridiagWork& t, const double* Z, size_t ldZ, size_t sZ, i64 m, E* C,
                   size_t ldC, size_t sC, void* work, hipStream_t s, bool first_only,
                   hipEvent_t promoted = nullptr) {
  const i64 n = t.n;
  const int K = t.K;
  E* V = (E*)work;
  E* G = V + (size_t)n * BT * K;
  E* T = G + (size_t)BT * BT * K;
  E* Y = T + (size_t)BT * BT * K;
  E* Y2 = Y + (size_t)BT * m * K;
  const size_t sV = (size_t)n * BT, sG = (size_t)BT * BT, sY = (size_t)BT * m;
  hipLaunchKernelGGL(k_bt_promote<E>, dim3(2048, K), dim3(256), 0, s, (size_t)n * m, Z, ldZ, sZ, n,
                     m, C, ldC, sC);
  FORGE_LAUNCH();
  if (promoted) FORGE_HIP(hipEventRecord(promoted, s));
  if (n < 2) return 0;
  // two-stage reduction: the chase's reflectors first (Q = Q1 Q2)
  if (t.two_stage && !first_only) CK_BT(twostage_apply_q2(t, m, C, ldC, sC, s));
  const int voff = t.two_stage ? t.voff : 1;
  const i64 nr = n - voff;  // reflectors of the (first-stage) reduction
  if (nr < 1) return 0;
  const Op CT = Op::H, NO = Op::N;
  const i64 nblocks = (nr + BT - 1) / BT;
  for (i64 bq = first_only ? 0 : nblocks - 1; bq >= 0; --bq) {
    const i64 i0 = bq * BT;
    const int nb = (int)((nr - i0) < BT ? (nr - i0) : BT);
    const i64 mt = n - i0 - voff;
    hipLaunchKernelGGL(k_bt_vclean<E>, dim3(cdiv(mt * nb, 256) < 4096 ? cdiv(mt * nb, 256) : 4096, K),
                       dim3(256), 0, s, n, i0, nb, voff, t.A, t.sA, t.lda, V, sV);
    FORGE_LAUNCH();
    // G = V^H V
    CK_BT(gemm_sb(CT, NO, nb, nb, mt, 1.0, V, n, sV, V, n, sV, 0.0, G, BT, sG, K, s));
    hipLaunchKernelGGL((k_bt_tbuild<E, BT>), dim3(nb, K), dim3(64), 0, s, nb, i0, n, G, t.tau, T);
    FORGE_LAUNCH();
    // Y = V^H C(i0+voff:n, :)
    CK_BT(gemm_sb(CT, NO, nb, m, mt, 1.0, V, n, sV, C + i0 + voff, ldC, sC, 0.0, Y, BT, sY, K, s));
    // Y2 = T Y
    CK_BT(gemm_sb(NO, NO, nb, m, nb, 1.0, T, BT, sG, Y, BT, sY, 0.0, Y2, BT, sY, K, s));
    // C(i0+voff:n, :) -= V Y2
    CK_BT(gemm_sb(NO, NO, mt, m, nb, -1.0, V, n, sV, Y2, BT, sY, 1.0, C + i0 + voff, ldC, sC, K, s));
  }
  return 0;
}

template <typename E>
static int apply_q_bt(int bt, const TridiagWork& t, const double* Z, size_t ldZ, size_t sZ, i64 m,
                      E* C, size_t ldC, size_t sC, void* work, hipStream_t s, bool first_only,
                      hipEvent_t promoted = nullptr) {
  switch (bt) {
    case 64: return apply_q<E, 64>(t, Z, ldZ, sZ, m, C, ldC, sC, work, s, first_only, promoted);
    case 128: return apply_q<E, 128>(t, Z, ldZ, sZ, m, C, ldC, sC, work, s, first_only, promoted);
    case 256: return apply_q<E, 256>(t, Z, ldZ, sZ, m, C, ldC, sC, work, s, first_only, promoted);
  }
  set_error("backtransform: unsupported block width");
  return -2;
}

static std::mutex g_bt_mu;
static std::map<std::tuple<int, i64, i64, int, int>, int> g_bt_width;  // (device, n, m, K, fp32)

// the tuned block width of a shape (measured on first use)
template <typename E>
static int tuned_width(const TridiagWork& t, const double* Z, size_t ldZ, size_t sZ, i64 m, E* C,
                       size_t ldC, size_t sC, void* work, hipStream_t s, int* width) {
  // FORGE_BT_WIDTH=64|128|256 in the environment forces the width (runs that
  // compare with a fixed width)
  if (const char* f = getenv("FORGE_BT_WIDTH")) {
    const int w = atoi(f);
    if (w == 64 || w == 128 || w == 256) {
      *width = w;
      return 0;
    }
  }
  int dev = 0;
  FORGE_HIP(hipGetDevice(&dev));
  const auto key = std::make_tuple(dev, t.n, m, t.K, (int)(sizeof(E) == sizeof(float2)));
  {
    std::lock_guard<std::mutex> g(g_bt_mu);
    auto it = g_bt_width.find(key);
    if (it != g_bt_width.end()) { *width = it->second; return 0; }
  }
  int best = 128;
  if (t.n > 64 && m > 0) {
    hipEvent_t e0, e1;
    FORGE_HIP(hipEventCreate(&e0));
    FORGE_HIP(hipEventCreate(&e1));
    double best_per = 0.0;
    for (int bt : BT_CAND) {
      float ms = 0.f;
      for (int rep = 0; rep < 2; ++rep) {
        int r = apply_q_bt<E>(bt, t, Z, ldZ, sZ, m, C, ldC, sC, work, s, true, e0);
        if (r) return r;
        FORGE_HIP(hipEventRecord(e1, s));
        FORGE_HIP(hipEventSynchronize(e1));
        FORGE_HIP(hipEventElapsedTime(&ms, e0, e1));
      }
      const double nb = (double)std::min<i64>(bt, t.n - (t.two_stage ? t.voff : 1));
      const double per = ms / nb;
      if (bt == BT_CAND[0] || per < best_per) { best_per = per; best = bt; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  std::lock_guard<std::mutex> g(g_bt_mu);
  g_bt_width[key] = best;
  *width = best;
  return 0;
}

int backtransform(const TridiagWork& t, const double* Z, size_t ldZ, size_t sZ, i64 m, z2* C,
                  size_t ldC, size_t sC, void* work, hipStream_t s, float2* C32, int* width_used,
                  int fixed_width) {
  int bt = 128;
  if (fixed_width) {
    // no tuning (it times with events of its own): the caller's width
    if (C32) { set_error("backtransform: a fixed width applies to the FP64 products only"); return -2; }
    if (width_used) *width_used = fixed_width;
    return apply_q_bt<z2>(fixed_width, t, Z, ldZ, sZ, m, C, ldC, sC, work, s, false);
  }
  if (!C32) {
    CK_BT(tuned_width<z2>(t, Z, ldZ, sZ, m, C, ldC, sC, work, s, &bt));
    if (width_used) *width_used = bt;
    return apply_q_bt<z2>(bt, t, Z, ldZ, sZ, m, C, ldC, sC, work, s, false);
  }
  const size_t ld32 = (size_t)t.n, s32 = (size_t)t.n * m;
  CK_BT(tuned_width<float2>(t, Z, ldZ, sZ, m, C32, ld32, s32, work, s, &bt));
  if (width_used) *width_used = bt;
  int r = apply_q_bt<float2>(bt, t, Z, ldZ, sZ, m, C32, ld32, s32, work, s, false);
  if (r) return r;
  if (m > 0) {
    hipLaunchKernelGGL(k_bt_normalise, dim3((unsigned)m, t.K), dim3(256), 0, s, t.n, C32, ld32, s32,
                       C, ldC, sC);
    FORGE_LAUNCH();
  }
  return 0;
}

}  // namespace forge
