// Matrix-core GEMMs of the eigensolver and the density products (gfx950).
//
//   zgemm / cgemm / dgemm   C = alpha op(A) op(B) + beta C, batched with
//                           strides; op = N or H (conjugate transpose; plain
//                           transpose for real data).  complex FP64, complex
//                           FP32 (the mixed-precision back-transformation) and
//                           real FP64 (the divide-and-conquer merges).
//   herk_cells              the bulk density of one k-point fused with its
//                           phased accumulation into the stored cells
//                           (Solve, HartreeFock.f90:456-475):
//                             R = C C^H (rank nocc, lower tiles only),
//                             rho_c(i,j) += ph_c conj(R(i,j)) for every
//                             (i, j), both triangles from the lower tiles;
//                           this replaces the reference's nLower-1 rank-1
//                           z_vector_mul updates (LapackRoutines.f90:168-183).
//   herk_lower              R = C C^H on the lower tiles (the C3-reduced set
//                           reads R at permuted positions).
//
// Complex products use four real products per complex multiply-add on the
// MFMA units (v_mfma_f64_16x16x4f64, v_mfma_f32_16x16x4f32):
//   Re += Ar Br - Ai Bi,  Im += Ar Bi + Ai Br,
// with the real and imaginary parts of each operand tile held as separate
// planes in LDS.  A block owns a BM x BN output tile; each wave owns a
// WM x WN sub-tile as (WM/16) x (WN/16) MFMA blocks of 16 x 16.  The K
// dimension advances in stages of 16: while a stage is multiplied from one
// LDS buffer, the next stage's operand tiles are loaded from global memory
// into registers and stored into the other buffer, so there is one barrier
// per stage.  Loads are coalesced along the contiguous dimension of each
// operand (rows for N, columns for H); a conjugate-transposed operand is
// transposed on its way into LDS, whose rows are padded against bank
// conflicts.
//
// Block order: consecutive blocks are dispatched round-robin over the 8
// XCDs, each with its own L2.  Block ids are remapped so every XCD works on
// a contiguous range of tiles, and tiles are ordered in groups of 8 tile
// rows, so the blocks resident on one XCD share operand panels in its L2.
#include "kernels.hpp"
#include "mfma.hpp"

namespace forge {

// storage element of a scalar type, real or complex
template <typename S, bool CPLX> struct Elem;
template <> struct Elem<double, true> { typedef z2 T; };
template <> struct Elem<float, true> { typedef float2 T; };
template <> struct Elem<double, false> { typedef double T; };
template <> struct Elem<float, false> { typedef float T; };

template <typename S> __device__ __forceinline__ void parts(const z2& v, S* p) { p[0] = v.x; p[1] = v.y; }
template <typename S> __device__ __forceinline__ void parts(const float2& v, S* p) { p[0] = v.x; p[1] = v.y; }
template <typename S> __device__ __forceinline__ void parts(const double& v, S* p) { p[0] = v; }
template <typename S> __device__ __forceinline__ void parts(const float& v, S* p) { p[0] = v; }
template <typename T> __device__ __forceinline__ T zero_el();
template <> __device__ __forceinline__ z2 zero_el<z2>() { return mk(0.0, 0.0); }
template <> __device__ __forceinline__ float2 zero_el<float2>() { return make_float2(0.f, 0.f); }
template <> __device__ __forceinline__ double zero_el<double>() { return 0.0; }
template <> __device__ __forceinline__ float zero_el<float>() { return 0.f; }

template <typename T> struct GemmArgs {
  i64 M, N, K;
  const T* A;
  i64 lda;
  size_t sA;
  const T* B;
  i64 ldb;
  size_t sB;
  T* C;
  i64 ldc;
  size_t sC;
  double alpha, beta;
  i64 tm, tn;  // tiles along M and N
  // herk_cells: nnc cell phases (complex) and the cell stride of rho
  const double* ph;
  int nnc;
  size_t cstride;
  int first;
};

enum { EPI_GEMM = 0, EPI_LOWER = 1, EPI_CELLS = 2 };

template <typename S, bool CPLX, bool TA, bool TB, int BM, int BN, int NWM, int NWN, int EPI>
__global__ void __launch_bounds__(NWM * NWN * 64)
    k_gemm(GemmArgs<typename Elem<S, CPLX>::T> g) {
  typedef typename Elem<S, CPLX>::T T;
  typedef Mfma<S> MF;
  typedef typename MF::acc ACC;
  constexpr int NT = NWM * NWN * 64;
  constexpr int BK = 16;
  constexpr int WM = BM / NWM, WN = BN / NWN, UM = WM / 16, UN = WN / 16;
  constexpr int NP = CPLX ? 2 : 1;
  constexpr int LDA_S = BM + 1, LDB_S = BN + 1;  // padded LDS rows
  constexpr int EA = BM * BK / NT, EB = BN * BK / NT;
  static_assert(BM * BK % NT == 0 && BN * BK % NT == 0, "tile / thread count");
  __shared__ S As[2][NP][BK][LDA_S];
  __shared__ S Bs[2][NP][BK][LDB_S];

  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int l16 = lane & 15, lk = lane >> 4;
  const int wm0 = (wv / NWN) * WM, wn0 = (wv % NWN) * WN;
  const i64 z = blockIdx.y;
  // tile of this block
  i64 I, J;
  if (EPI == EPI_GEMM) {
    const i64 nb = g.tm * g.tn;
    const i64 pid = xcd_remap(blockIdx.x, nb);
    constexpr i64 GM = 8;
    const i64 grp = pid / (GM * g.tn), first_m = grp * GM;
    const i64 gm = (g.tm - first_m) < GM ? (g.tm - first_m) : GM;
    I = first_m + (pid % (GM * g.tn)) % gm;
    J = (pid % (GM * g.tn)) / gm;
  } else {
    const i64 nb = g.tm * (g.tm + 1) / 2;
    const i64 t = xcd_remap(blockIdx.x, nb);
    I = (i64)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= t) ++I;
    while (I * (I + 1) / 2 > t) --I;
    J = t - I * (I + 1) / 2;
  }
  const i64 m0 = I * BM, n0 = J * BN;
  const T* A = g.A + z * g.sA;
  const T* B = g.B + z * g.sB;
  const i64 M = g.M, N = g.N, K = g.K;

  ACC re[UM][UN], im[UM][UN];
#pragma unroll
  for (int u = 0; u < UM; ++u)
#pragma unroll
    for (int v = 0; v < UN; ++v) {
      re[u][v] = ACC{};
      im[u][v] = ACC{};
    }

  T ra[EA], rb[EB];
  auto load = [&](i64 k0) {
#pragma unroll
    for (int q = 0; q < EA; ++q) {
      const int idx = tid + NT * q;
      int mm, kk;
      if (!TA) { mm = idx % BM; kk = idx / BM; }
      else { kk = idx % BK; mm = idx / BK; }
      const i64 gm_ = m0 + mm, gk = k0 + kk;
      ra[q] = (gm_ < M && gk < K) ? (TA ? A[gk + g.lda * gm_] : A[gm_ + g.lda * gk]) : zero_el<T>();
    }
#pragma unroll
    for (int q = 0; q < EB; ++q) {
      const int idx = tid + NT * q;
      int nn, kk;
      if (!TB) { kk = idx % BK; nn = idx / BK; }
      else { nn = idx % BN; kk = idx / BN; }
      const i64 gn = n0 + nn, gk = k0 + kk;
      rb[q] = (gn < N && gk < K) ? (TB ? B[gn + g.ldb * gk] : B[gk + g.ldb * gn]) : zero_el<T>();
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < EA; ++q) {
      const int idx = tid + NT * q;
      int mm, kk;
      if (!TA) { mm = idx % BM; kk = idx / BM; }
      else { kk = idx % BK; mm = idx / BK; }
      S p[2];
      parts<S>(ra[q], p);
      As[buf][0][kk][mm] = p[0];
      if (CPLX) As[buf][NP - 1][kk][mm] = TA ? -p[1] : p[1];
    }
#pragma unroll
    for (int q = 0; q < EB; ++q) {
      const int idx = tid + NT * q;
      int nn, kk;
      if (!TB) { kk = idx % BK; nn = idx / BK; }
      else { nn = idx % BN; kk = idx / BN; }
      S p[2];
      parts<S>(rb[q], p);
      Bs[buf][0][kk][nn] = p[0];
      if (CPLX) Bs[buf][NP - 1][kk][nn] = TB ? -p[1] : p[1];
    }
  };

  const i64 nst = (K + BK - 1) / BK;
  if (nst > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (i64 st = 0; st < nst; ++st) {
    const int buf = (int)(st & 1);
    if (st + 1 < nst) load((st + 1) * BK);
#pragma unroll
    for (int kq = 0; kq < BK; kq += 4) {
      const int kr = kq + lk;
      S a[UM][NP], b[UN][NP];
#pragma unroll
      for (int u = 0; u < UM; ++u)
#pragma unroll
        for (int p = 0; p < NP; ++p) a[u][p] = As[buf][p][kr][wm0 + u * 16 + l16];
#pragma unroll
      for (int v = 0; v < UN; ++v)
#pragma unroll
        for (int p = 0; p < NP; ++p) b[v][p] = Bs[buf][p][kr][wn0 + v * 16 + l16];
#pragma unroll
      for (int u = 0; u < UM; ++u)
#pragma unroll
        for (int v = 0; v < UN; ++v) {
          re[u][v] = MF::mma(a[u][0], b[v][0], re[u][v]);
          if (CPLX) {
            re[u][v] = MF::mma(a[u][NP - 1], -b[v][NP - 1], re[u][v]);
            im[u][v] = MF::mma(a[u][0], b[v][NP - 1], im[u][v]);
            im[u][v] = MF::mma(a[u][NP - 1], b[v][0], im[u][v]);
          }
        }
    }
    if (st + 1 < nst) store(buf ^ 1);
    __syncthreads();
  }

  // epilogue
  T* C = g.C + z * g.sC;
#pragma unroll
  for (int u = 0; u < UM; ++u)
#pragma unroll
    for (int v = 0; v < UN; ++v)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const i64 r = m0 + wm0 + u * 16 + MF::row(lane, i);
        const i64 c = n0 + wn0 + v * 16 + l16;
        if (r >= M || c >= N) continue;
        if constexpr (EPI == EPI_CELLS) {
          // R(r, c) = re + i im; rho_q(r, c) += ph_q conj(R), rho_q(c, r) += ph_q R
          const double x = (double)re[u][v][i], y = CPLX ? (double)im[u][v][i] : 0.0;
          z2* rho = (z2*)C;
          const bool diag = I == J;
          for (int q = 0; q < g.nnc; ++q) {
            const double px = g.ph[2 * q], py = g.ph[2 * q + 1];
            z2* cell = rho + (size_t)q * g.cstride;
            const z2 lo = mk(px * x + py * y, py * x - px * y);  // ph conj(R)
            const size_t o1 = (size_t)r + (size_t)g.ldc * c;
            cell[o1] = g.first ? lo : cell[o1] + lo;
            if (!diag) {
              const z2 up = mk(px * x - py * y, py * x + px * y);  // ph R
              const size_t o2 = (size_t)c + (size_t)g.ldc * r;
              cell[o2] = g.first ? up : cell[o2] + up;
            }
          }
        } else {
          const size_t o = (size_t)r + (size_t)g.ldc * c;
          if constexpr (CPLX) {
            typedef decltype(C[0].x) SS;
            T val;
            val.x = (SS)(g.alpha * (double)re[u][v][i]);
            val.y = (SS)(g.alpha * (double)im[u][v][i]);
            if (g.beta != 0.0) {
              const T old = C[o];
              val.x = (SS)((double)val.x + g.beta * (double)old.x);
              val.y = (SS)((double)val.y + g.beta * (double)old.y);
            }
            C[o] = val;
          } else {
            double val = g.alpha * (double)re[u][v][i];
            if (g.beta != 0.0) val += g.beta * (double)C[o];
            C[o] = (T)val;
          }
        }
      }
}

// -------------------------------------------------------------------- host side

template <typename S, bool CPLX, bool TA, bool TB, int BM, int BN, int NWM, int NWN, int EPI>
static int launch(const GemmArgs<typename Elem<S, CPLX>::T>& g0, i64 tiles, int batch,
                  hipStream_t s) {
  if (tiles <= 0 || batch <= 0) return 0;
  if (tiles > 0x7fffffff || batch > 65535) {
    set_error("gemm: grid too large");
    return -2;
  }
  hipLaunchKernelGGL((k_gemm<S, CPLX, TA, TB, BM, BN, NWM, NWN, EPI>), dim3((unsigned)tiles, batch),
                     dim3(NWM * NWN * 64), 0, s, g0);
  FORGE_LAUNCH();
  return 0;
}

static bool bad_dims(i64 M, i64 N, i64 K, i64 lda, i64 ldb, i64 ldc, Op ta, Op tb) {
  if (M < 0 || N < 0 || K < 0) return true;
  const i64 ra = ta == Op::N ? M : K, rb = tb == Op::N ? K : N;
  return lda < (ra > 1 ? ra : 1) || ldb < (rb > 1 ? rb : 1) || ldc < (M > 1 ? M : 1);
}

// large outputs: 128 x 128 tiles, 8 waves (64 x 32 each); small outputs
// (fewer than about two tiles per CU): 64 x 64 tiles, 4 waves (32 x 32)
template <typename S, bool CPLX>
static int gemm_dispatch(Op ta, Op tb, const GemmArgs<typename Elem<S, CPLX>::T>& g0, int batch,
                         hipStream_t s) {
  GemmArgs<typename Elem<S, CPLX>::T> g = g0;
  const bool big = (double)g.M * (double)g.N * batch >= 512.0 * 128 * 128;
#define GEMM_CASE(TA_, TB_)                                                                  \
  if ((ta == Op::H) == TA_ && (tb == Op::H) == TB_) {                                        \
    if (big) {                                                                               \
      g.tm = cdiv(g.M, 128);                                                                 \
      g.tn = cdiv(g.N, 128);                                                                 \
      return launch<S, CPLX, TA_, TB_, 128, 128, 2, 4, EPI_GEMM>(g, g.tm * g.tn, batch, s);  \
    }                                                                                        \
    g.tm = cdiv(g.M, 64);                                                                    \
    g.tn = cdiv(g.N, 64);                                                                    \
    return launch<S, CPLX, TA_, TB_, 64, 64, 2, 2, EPI_GEMM>(g, g.tm * g.tn, batch, s);      \
  }
  GEMM_CASE(false, false)
  GEMM_CASE(true, false)
  GEMM_CASE(false, true)
  GEMM_CASE(true, true)
#undef GEMM_CASE
  return -2;
}

template <typename T>
static GemmArgs<T> args(i64 M, i64 N, i64 K, double alpha, const T* A, i64 lda, size_t sA,
                        const T* B, i64 ldb, size_t sB, double beta, T* C, i64 ldc, size_t sC) {
  GemmArgs<T> g{};
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.sA = sA;
  g.B = B; g.ldb = ldb; g.sB = sB;
  g.C = C; g.ldc = ldc; g.sC = sC;
  g.alpha = alpha; g.beta = beta;
  return g;
}

int zgemm(Op ta, Op tb, i64 M, i64 N, i64 K, double alpha, const z2* A, i64 lda, size_t sA,
          const z2* B, i64 ldb, size_t sB, double beta, z2* C, i64 ldc, size_t sC, int batch,
          hipStream_t s) {
  if (bad_dims(M, N, K, lda, ldb, ldc, ta, tb)) { set_error("zgemm: invalid dimensions"); return -2; }
  if (M == 0 || N == 0) return 0;
  return gemm_dispatch<double, true>(ta, tb, args(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC),
                                     batch, s);
}

int cgemm(Op ta, Op tb, i64 M, i64 N, i64 K, double alpha, const float2* A, i64 lda, size_t sA,
          const float2* B, i64 ldb, size_t sB, double beta, float2* C, i64 ldc, size_t sC,
          int batch, hipStream_t s) {
  if (bad_dims(M, N, K, lda, ldb, ldc, ta, tb)) { set_error("cgemm: invalid dimensions"); return -2; }
  if (M == 0 || N == 0) return 0;
  return gemm_dispatch<float, true>(ta, tb, args(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC),
                                    batch, s);
}

int dgemm(Op ta, Op tb, i64 M, i64 N, i64 K, double alpha, const double* A, i64 lda, size_t sA,
          const double* B, i64 ldb, size_t sB, double beta, double* C, i64 ldc, size_t sC,
          int batch, hipStream_t s) {
  if (bad_dims(M, N, K, lda, ldb, ldc, ta, tb)) { set_error("dgemm: invalid dimensions"); return -2; }
  if (M == 0 || N == 0) return 0;
  return gemm_dispatch<double, false>(ta, tb, args(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC),
                                      batch, s);
}

int herk_lower(i64 n, i64 k, const z2* A, i64 lda, z2* R, i64 ldr, hipStream_t s) {
  if (n < 0 || k < 0 || lda < (n > 1 ? n : 1) || ldr < (n > 1 ? n : 1)) {
    set_error("herk_lower: invalid dimensions");
    return -2;
  }
  if (n == 0) return 0;
  GemmArgs<z2> g = args<z2>(n, n, k, 1.0, A, lda, 0, A, lda, 0, 0.0, R, ldr, 0);
  g.tm = g.tn = cdiv(n, 128);
  return launch<double, true, false, true, 128, 128, 2, 4, EPI_LOWER>(g, g.tm * (g.tm + 1) / 2, 1, s);
}

int herk_cells(i64 n, i64 k, const z2* A, i64 lda, int nnc, const double* ph, z2* rho,
               bool first, hipStream_t s) {
  if (n < 0 || k < 0 || lda < (n > 1 ? n : 1) || nnc < 1) {
    set_error("herk_cells: invalid dimensions");
    return -2;
  }
  if (n == 0) return 0;
  GemmArgs<z2> g = args<z2>(n, n, k, 1.0, A, lda, 0, A, lda, 0, 0.0, rho, n, 0);
  g.tm = g.tn = cdiv(n, 128);
  g.ph = ph;
  g.nnc = nnc;
  g.cstride = (size_t)n * n;
  g.first = first ? 1 : 0;
  return launch<double, true, false, true, 128, 128, 2, 4, EPI_CELLS>(g, g.tm * (g.tm + 1) / 2, 1, s);
}

}  // namespace forge
