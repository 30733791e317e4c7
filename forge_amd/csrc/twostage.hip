// Two-stage reduction of K Hermitian matrices (lower triangle, column-major)
// to real symmetric tridiagonal form, A = Q1 Q2 T Q2^H Q1^H, for the
// eigensolver that replaces zheevr (LapackRoutines.f90:411-428).  The
// one-stage reduction (tridiag.hip) streams the whole trailing matrix once
// per column for its Householder vector (n matrix-vector products bound by
// HBM bandwidth); here the trailing matrix is read once per panel of TS_KD
// columns, by matrix-core products.
//
// Stage 1, dense -> band of width KD = TS_KD, per panel of columns i0 .. i0+KD-1
// (rows r0 = i0 + KD .. n-1 below the panel's diagonal block, mp = n - r0):
//   panel QR      A(r0:n, i0:i0+KD) = Q R, Q = H_0 ... H_{nq-1} = I - V T V^H
//                 (zgeqr2 column by column: k_ts_qa forms the reflector and the
//                 partial sums of v^H A over the remaining columns, k_ts_qb
//                 applies them and forms the next column's norm partials);
//                 R stays in the band, the reflector tails below it
//   T             zlarft 'F','C' from V^H V (k_ts_tbuild)
//   Y = A22 V     the trailing matrix's lower triangle read as a Hermitian
//                 matrix on the FP64 matrix cores, split over columns (k_ts_hemm)
//   X = Y T, W = X - V (T^H V^H X) / 2     (k_ts_x, k_ts_mz, k_ts_w)
//   A22 -= V W^H + W V^H                    (k_trd_her2k_mfma, tridiag.hip)
// which is A22 <- Q^H A22 Q.
//
// Stage 2, band -> tridiagonal (bulge chasing in band storage AB, column c
// holding A(c .. c + 2 KD, c): the band plus room for the bulge).  Sweep s
// annihilates column s below its subdiagonal; its step t (t = 0, 1, ...)
// applies the reflector H_{s,t} on rows st = s + 1 + t KD .. st + KD - 1 from
// both sides to the diagonal block there, then from the right to the block
// below (rows st + KD ..), whose first column then gives H_{s,t+1}, applied
// from the left to that block's other columns.  Step (s, t) only depends on
// (s, t-1) and (s-1, t+1), and the steps with 2 s + t = slot touch disjoint
// parts of the band, so one launch per slot runs all of them (k_ts_chase).
//
// The chase's reflectors are stored by groups of NB = TS_NB consecutive
// sweeps and step t: block (g, t) holds H_{s,t} for s = g NB .. g NB + NB - 1,
// each on rows g NB + 1 + t KD + i .. + KD - 1 (a parallelogram of KD + NB - 1
// rows).  Q2 is applied to the eigenvectors group by group, last group first,
// and inside a group step by step, t = 0, 1, ...: every block as
// C <- C - V T (V^H C) with its own triangular factor T (k_ts_t2, k_ts_q2).
// That order respects every overlap of rows between the reflectors.
#include <hip/hip_ext.h>

#include <algorithm>

#include "kernels.hpp"
#include "mfma.hpp"

namespace forge {

constexpr int KD = TS_KD;
constexpr int NBG = TS_NB;
constexpr int LDAB = TS_LDAB;
constexpr int QROWS = 256;  // rows per block of the panel QR kernels
static_assert(KD == 32 && NBG == 32, "the kernels below assume 32-wide panels and groups");

__host__ __device__ inline i64 nbq_of(i64 mp) { return (mp + QROWS - 1) / QROWS; }
__host__ __device__ inline i64 nbx_of(i64 mp) { return (mp + 63) / 64; }
// chase steps of sweep s (0 <= s <= n - 2)
__host__ __device__ inline i64 ts_nsteps(i64 n, i64 s) { return (n - 2 - s) / KD + 1; }
static i64 groups_of(i64 n) { return n >= 2 ? (n - 1 + NBG - 1) / NBG : 0; }

struct TsLayout {
  size_t V, X, Y, qp, gp, np, T, Zm, mp, AB, V2, T2, tau2, goff, total;
  i64 G, nblk2;
};

static TsLayout ts_layout(i64 n, int K) {
  TsLayout L{};
  L.G = groups_of(n);
  L.nblk2 = 0;
  for (i64 g = 0; g < L.G; ++g) L.nblk2 += ts_nsteps(n, g * NBG);
  const i64 nbq = nbq_of(n), nbx = nbx_of(n);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t Kz = (size_t)K * sizeof(z2);
  L.V = take(Kz * n * KD);
  L.X = take(Kz * n * KD);
  L.Y = take(Kz * TS_YSPLIT * n * KD);
  L.qp = take(Kz * nbq * KD);
  L.gp = take(Kz * nbq * KD * KD);
  L.np = take(sizeof(double) * (size_t)K * nbq);
  L.T = take(Kz * KD * KD);
  L.Zm = take(Kz * KD * KD);
  L.mp = take(Kz * nbx * KD * KD);
  L.AB = take(Kz * n * LDAB);
  L.V2 = take(Kz * L.nblk2 * NBG * KD);
  L.T2 = take(Kz * L.nblk2 * NBG * NBG);
  L.tau2 = take(Kz * L.nblk2 * NBG);
  L.goff = take(sizeof(i64) * (L.G + 1));
  L.total = off;
  return L;
}

size_t twostage_workspace_bytes(i64 n, int K) { return ts_layout(n, K).total; }

int twostage_bind(TwoStageWork& ts, void* base, i64 n, int K) {
  const TsLayout L = ts_layout(n, K);
  char* b = (char*)base;
  ts.n = n;
  ts.K = K;
  ts.V = (z2*)(b + L.V);
  ts.X = (z2*)(b + L.X);
  ts.Y = (z2*)(b + L.Y);
  ts.qpart = (z2*)(b + L.qp);
  ts.gpart = (z2*)(b + L.gp);
  ts.npart = (double*)(b + L.np);
  ts.T = (z2*)(b + L.T);
  ts.Zm = (z2*)(b + L.Zm);
  ts.mpart = (z2*)(b + L.mp);
  ts.AB = (z2*)(b + L.AB);
  ts.V2 = (z2*)(b + L.V2);
  ts.T2 = (z2*)(b + L.T2);
  ts.tau2 = (z2*)(b + L.tau2);
  ts.goff = (i64*)(b + L.goff);
  ts.G = L.G;
  ts.nblk2 = L.nblk2;
  ts.hoff.assign((size_t)L.G + 1, 0);
  for (i64 g = 0; g < L.G; ++g) ts.hoff[g + 1] = ts.hoff[g] + ts_nsteps(n, g * NBG);
  FORGE_HIP(hipMemcpy(ts.goff, ts.hoff.data(), sizeof(i64) * (L.G + 1), hipMemcpyHostToDevice));
  return 0;
}

// ---------------------------------------------------------------- helpers
struct TsRefl {
  z2 tau, scale;
  double beta;
};

// zlarfg: H = I - tau v v^H, v = (1, x / (alpha - beta)), H^H (alpha, x) = (beta, 0)
__device__ inline TsRefl ts_reflector(z2 alpha, double xn2) {
  TsRefl R;
  if (xn2 == 0.0 && alpha.y == 0.0) {
    R.tau = mk(0.0, 0.0);
    R.beta = alpha.x;
    R.scale = mk(0.0, 0.0);
  } else {
    const double nr = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
    R.beta = alpha.x >= 0.0 ? -nr : nr;
    R.tau = mk((R.beta - alpha.x) / R.beta, -alpha.y / R.beta);
    const z2 den = mk(alpha.x - R.beta, alpha.y);
    const double dd = abs2(den);
    R.scale = mk(den.x / dd, -den.y / dd);
  }
  return R;
}

__device__ inline double ts_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline double ts_block_sum256(double v, double* sh) {
  v = ts_wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

// Loads below are issued unconditionally -- from `safe`, an element known to
// hold a finite value, when the wanted one does not exist -- and the value is
// scaled by 1 or 0 afterwards.  A load behind a per-element branch (or a
// select the compiler turns back into one) makes it wait for each load before
// issuing the next.
__device__ inline z2 ld_or_zero(const z2* p, size_t off, bool ok, size_t safe) {
  const z2 v = p[ok ? off : safe];
  const double f = ok ? 1.0 : 0.0;
  return mk(f * v.x, f * v.y);
}

// A22(i, j) of the trailing matrix at r0 (local indices), from the lower
// triangle (mirrored and conjugated above the diagonal); zero unless ok
__device__ inline z2 herm_at(const z2* A, i64 lda, i64 r0, i64 i, i64 j, bool ok) {
  const bool up = i < j;
  const i64 a = up ? j : i, b = up ? i : j;
  z2 v = ld_or_zero(A, (size_t)((r0 + a) + lda * (r0 + b)), ok, (size_t)(r0 + lda * r0));
  if (up) v.y = -v.y;
  if (i == j) v.y = 0.0;
  return v;
}

// ---------------------------------------------------------------- panel QR
// norm partials of the panel's first column below its first panel row
__global__ void __launch_bounds__(256) k_ts_qnorm(i64 n, i64 i0, const z2* A_, size_t sA, i64 lda,
                                                  double* npart_, i64 nbqmax) {
  const int k = blockIdx.y;
  const z2* A = A_ + k * sA;
  const i64 r0 = i0 + KD, mp = n - r0;
  const i64 rr = (i64)blockIdx.x * QROWS + threadIdx.x;
  __shared__ double sh[4];
  double v = abs2(ld_or_zero(A, (size_t)((r0 + rr) + lda * i0), rr >= 1 && rr < mp,
                             (size_t)(r0 + lda * i0)));
  v = ts_block_sum256(v, sh);
  if (threadIdx.x == 0) npart_[(size_t)k * nbqmax + blockIdx.x] = v;
}

// Reflector of panel column jj from the norm partials and the pivot (read by
// every block, so left unchanged here: k_ts_qb writes beta to it); v written
// to V and its tail to A; partial sums per block of
//   w_q = v^H A(:, i0 + q)  (q > jj)   and   G(q, jj) = V(:, q)^H v  (q < jj).
__global__ void __launch_bounds__(256) k_ts_qa(i64 n, i64 i0, int jj, int nq, z2* A_, size_t sA,
                                               i64 lda, z2* V_, z2* tau_, const double* npart_,
                                               z2* qpart_, z2* gpart_, i64 nbqmax) {
  const int k = blockIdx.y;
  z2* A = A_ + k * sA;
  z2* V = V_ + (size_t)k * n * KD;
  const i64 r0 = i0 + KD, mp = n - r0, nbq = nbq_of(mp), c = i0 + jj;
  __shared__ TsRefl Rs;
  __shared__ z2 red[4][KD];
  if (threadIdx.x < 64) {
    double s = 0.0;
    for (i64 b = threadIdx.x; b < nbq; b += 64) s += npart_[(size_t)k * nbqmax + b];
    s = ts_wave_sum(s);
    if (threadIdx.x == 0) Rs = ts_reflector(A[(r0 + jj) + lda * c], s);
  }
  __syncthreads();
  const TsRefl R = Rs;
  const i64 rr = (i64)blockIdx.x * QROWS + threadIdx.x, r = r0 + rr;
  const bool live = rr < mp;
  const i64 rs = live ? r : r0;  // a row inside the matrix for the unconditional loads
  z2 v = ld_or_zero(A, (size_t)(rs + lda * c), live && rr > jj, (size_t)(r0 + lda * c)) * R.scale;
  if (live && rr == jj) v = mk(1.0, 0.0);
  if (live) {
    V[r + n * jj] = v;
    if (rr > jj) A[r + lda * c] = v;
    if (rr == jj) {
      // beta goes to the pivot in k_ts_qb: every block of this launch reads
      // the pivot for its reflector, so it must not change before they have;
      // it is parked in slot 0 of block 0's partials (qpart holds q > jj only)
      qpart_[((size_t)k * nbqmax) * KD] = mk(R.beta, 0.0);
      tau_[(size_t)k * n + c] = R.tau;
    }
    if (jj == 0)
      for (int q = nq; q < KD; ++q) V[r + n * q] = mk(0.0, 0.0);
  }
  if (jj == 0 && blockIdx.x == 0 && threadIdx.x >= nq && threadIdx.x < KD)
    tau_[(size_t)k * n + i0 + threadIdx.x] = mk(0.0, 0.0);
  // slot q: the Gram entry for q < jj, the product with column q for q > jj.
  // All 32 loads are issued first; the 32 sums over the wave's 64 lanes are
  // then formed by a transposing butterfly: at each of the offsets 32, 16, 8,
  // 4, 2 a lane keeps half of its slots and adds its partner's copy of them
  // (16 + 8 + 4 + 2 + 1 exchanges), a last exchange at offset 1 completes
  // the sum, and lane l holds slot l / 2.
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  z2 pq[KD];
#pragma unroll
  for (int q = 0; q < KD; ++q) {
    // q < jj: V(r, q); q > jj: A(r, i0 + q) (v is zero on rows above jj)
    const z2* src = q < jj ? V + (rs + n * q) : A + (rs + lda * (i0 + q));
    pq[q] = *src;
  }
#pragma unroll
  for (int q = 0; q < KD; ++q) {
    z2 p = q < jj ? cmul(pq[q], v) : cmul(v, pq[q]);
    if (!live || q == jj) p = mk(0.0, 0.0);
    pq[q] = p;
  }
#pragma unroll
  for (int half = KD / 2; half >= 1; half >>= 1) {
    const bool up = (lane & (2 * half)) != 0;  // offsets 32, 16, 8, 4, 2
#pragma unroll
    for (int j = 0; j < half; ++j) {
      const z2 keep = up ? pq[j + half] : pq[j];
      const z2 send = up ? pq[j] : pq[j + half];
      pq[j] = mk(keep.x + __shfl_xor(send.x, 2 * half, 64), keep.y + __shfl_xor(send.y, 2 * half, 64));
    }
  }
  {
    const z2 tot = mk(pq[0].x + __shfl_xor(pq[0].x, 1, 64), pq[0].y + __shfl_xor(pq[0].y, 1, 64));
    if ((lane & 1) == 0) red[wv][lane >> 1] = tot;
  }
  __syncthreads();
  if (threadIdx.x < KD) {
    const int q = threadIdx.x;
    const z2 s = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
    if (q > jj) qpart_[((size_t)k * nbqmax + blockIdx.x) * KD + q] = s;
    else if (q < jj) gpart_[(((size_t)k * nbqmax + blockIdx.x) * KD + jj) * KD + q] = s;
  }
}

// A(:, i0 + q) -= conj(tau) v w_q for q > jj (own rows), beta to the pivot
// of column jj, norm partials of column jj + 1 below its first panel row
__global__ void __launch_bounds__(256) k_ts_qb(i64 n, i64 i0, int jj, int nq, z2* A_, size_t sA,
                                               i64 lda, const z2* V_, const z2* tau_,
                                               const z2* qpart_, double* npart_, i64 nbqmax) {
  const int k = blockIdx.y;
  z2* A = A_ + k * sA;
  const z2* V = V_ + (size_t)k * n * KD;
  const i64 r0 = i0 + KD, mp = n - r0, nbq = nbq_of(mp), c = i0 + jj;
  __shared__ z2 part[8][KD];
  __shared__ z2 wl[KD];
  __shared__ double sh[4];
  {
    const int q = threadIdx.x % KD, ch = threadIdx.x / KD;
    z2 s = mk(0.0, 0.0);
    if (q > jj)
      for (i64 b = ch; b < nbq; b += 8) s = s + qpart_[((size_t)k * nbqmax + b) * KD + q];
    part[ch][q] = s;
  }
  __syncthreads();
  if (threadIdx.x < KD) {
    z2 s = part[0][threadIdx.x];
#pragma unroll
    for (int h = 1; h < 8; ++h) s = s + part[h][threadIdx.x];
    wl[threadIdx.x] = s;
  }
  __syncthreads();
  const z2 tb = conjz(tau_[(size_t)k * n + c]);
  const i64 rr = (i64)blockIdx.x * QROWS + threadIdx.x, r = r0 + rr;
  double nrm = 0.0;
  const bool act = rr < mp && rr >= jj;
  const i64 rs = act ? r : r0 + jj;
  const z2 tv = tb * V[rs + n * jj];
  z2 col[KD];
#pragma unroll
  for (int q = 0; q < KD; ++q) col[q] = A[rs + lda * (i0 + (q > jj ? q : jj))];
  if (act) {
#pragma unroll
    for (int q = 0; q < KD; ++q)
      if (q > jj) {
        const z2 a = col[q] - tv * wl[q];
        A[r + lda * (i0 + q)] = a;
        if (q == jj + 1 && rr >= jj + 2) nrm = abs2(a);
      }
    // the pivot of column jj becomes beta (k_ts_qa parked it); no other
    // thread of this launch reads it
    if (rr == jj) A[r + lda * c] = qpart_[((size_t)k * nbqmax) * KD];
  }
  if (jj + 1 < nq) {
    nrm = ts_block_sum256(nrm, sh);
    if (threadIdx.x == 0) npart_[(size_t)k * nbqmax + blockIdx.x] = nrm;
  }
}

// T (KD x KD upper, zlarft 'F','C'): T(j, j) = tau_j,
// T(0:j, j) = -tau_j T(0:j, 0:j) G(0:j, j), G = V^H V; zero for j >= nq
__global__ void __launch_bounds__(256) k_ts_tbuild(i64 n, i64 i0, int nq, const z2* tau_,
                                                   const z2* gpart_, i64 nbq, i64 nbqmax, z2* T_) {
  const int k = blockIdx.x;
  __shared__ z2 Gs[KD][KD + 1];
  __shared__ z2 Ts[KD][KD + 1];
  __shared__ z2 tl[KD];
  for (int e = threadIdx.x; e < KD * KD; e += 256) {
    const int i = e % KD, j = e / KD;
    z2 s = mk(0.0, 0.0);
    if (i < j && j < nq)
      for (i64 b = 0; b < nbq; ++b) s = s + gpart_[(((size_t)k * nbqmax + b) * KD + j) * KD + i];
    Gs[i][j] = s;
  }
  if (threadIdx.x < KD)
    tl[threadIdx.x] = threadIdx.x < nq ? tau_[(size_t)k * n + i0 + threadIdx.x] : mk(0.0, 0.0);
  __syncthreads();
  const int i = threadIdx.x;
  for (int j = 0; j < KD; ++j) {
    if (i < j) {
      z2 s = mk(0.0, 0.0);
      for (int l = i; l < j; ++l) s = s + Ts[i][l] * Gs[l][j];
      Ts[i][j] = (-1.0) * (tl[j] * s);
    } else if (i == j) {
      Ts[j][j] = tl[j];
    } else if (i < KD) {
      Ts[i][j] = mk(0.0, 0.0);
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < KD * KD; e += 256) T_[(size_t)k * KD * KD + e] = Ts[e % KD][e / KD];
}

// ---------------------------------------------------------------- Y = A22 V
// Block: 64 rows of Y (tile) over the column range [kb, ke) of A22 (split q).
// Wave w owns rows 16 w .. 16 w + 15 and all 32 columns of V: two 16 x 16
// FP64 MFMA blocks (complex, three real products each).  The A operand (one
// element per lane per k-step: row lane % 16, column lane / 16) is read
// straight from the lower triangle (mirrored and conjugated above the
// diagonal), a stage of HS columns ahead of its use; V's stage of HS rows goes
// through LDS, double-buffered.  HS = 32: eight loads per lane in flight
// across each barrier.
constexpr int HS = 32;
__global__ void __launch_bounds__(256) k_ts_hemm(i64 n, i64 r0, i64 mp, int nq, const z2* A_,
                                                 size_t sA, i64 lda, const z2* V_, z2* Y_,
                                                 i64 ntile, i64 KS) {
  typedef Mfma<double> MF;
  typedef CAcc<double, true> CA;
  const int k = blockIdx.y;
  const z2* A = A_ + k * sA;
  const z2* V = V_ + (size_t)k * n * KD;
  const i64 tile = blockIdx.x % ntile, q = blockIdx.x / ntile;
  const i64 kb = q * KS, ke = std::min<i64>(mp, kb + KS);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, l16 = lane & 15, lk = lane >> 4;
  const i64 row = 64 * tile + 16 * wv + l16;
  __shared__ double Vs[2][2][HS][KD + 1];
  CA acc[2];
  acc[0].zero();
  acc[1].zero();
  constexpr int NKS = HS / 4, NV = HS * KD / 256;
  z2 af[NKS], vf[NV];
  const i64 nst = (ke - kb + HS - 1) / HS;
  auto load = [&](i64 st) {
    const i64 c0 = kb + HS * st;
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      const i64 col = c0 + 4 * ks + lk;
      af[ks] = herm_at(A, lda, r0, row, col, row < mp && col < ke);
    }
#pragma unroll
    for (int h = 0; h < NV; ++h) {
      const int e = tid + 256 * h, vr = e % HS, vc = e / HS;
      const i64 gr = c0 + vr;
      vf[h] = ld_or_zero(V, (size_t)((r0 + gr) + n * vc), gr < ke && vc < nq, (size_t)r0);
    }
  };
  auto store_v = [&](int buf) {
#pragma unroll
    for (int h = 0; h < NV; ++h) {
      const int e = tid + 256 * h, vr = e % HS, vc = e / HS;
      Vs[buf][0][vr][vc] = vf[h].x;
      Vs[buf][1][vr][vc] = vf[h].y;
    }
  };
  if (nst > 0) {
    load(0);
    store_v(0);
  }
  __syncthreads();
  for (i64 st = 0; st < nst; ++st) {
    const int buf = (int)(st & 1);
    z2 cur[NKS];
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) cur[ks] = af[ks];
    if (st + 1 < nst) load(st + 1);
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      const int kk = 4 * ks + lk;
#pragma unroll
      for (int cb = 0; cb < 2; ++cb)
        acc[cb].mma(cur[ks].x, cur[ks].y, Vs[buf][0][kk][16 * cb + l16], Vs[buf][1][kk][16 * cb + l16]);
    }
    if (st + 1 < nst) store_v(buf ^ 1);
    __syncthreads();
  }
  z2* Y = Y_ + ((size_t)k * TS_YSPLIT + q) * n * KD;
#pragma unroll
  for (int cb = 0; cb < 2; ++cb)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const i64 r = 64 * tile + 16 * wv + MF::row(lane, i);
      if (r < mp) Y[(r0 + r) + n * (16 * cb + l16)] = mk(acc[cb].re(i), acc[cb].im(i));
    }
}

// X = (sum of the splits of Y) T for 64 rows, and the rows' partial V^H X
__global__ void __launch_bounds__(256) k_ts_x(i64 n, i64 r0, i64 mp, int nsplit, const z2* Y_,
                                              const z2* T_, const z2* V_, z2* X_, z2* mpart_,
                                              i64 nbxmax) {
  const int k = blockIdx.y;
  const i64 b = blockIdx.x;
  const int lr = threadIdx.x / 4, cq = threadIdx.x % 4;
  const i64 r = 64 * b + lr;
  const bool live = r < mp;
  __shared__ z2 Ts[KD][KD + 1];
  __shared__ z2 Bf[64][KD + 1];
  __shared__ z2 Xs[64][KD + 1];
  for (int e = threadIdx.x; e < KD * KD; e += 256) Ts[e % KD][e / KD] = T_[(size_t)k * KD * KD + e];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int col = 8 * cq + c;
    z2 y = mk(0.0, 0.0);
    const i64 rs = live ? r : 0;
    for (int q = 0; q < nsplit; ++q)
      y = y + Y_[((size_t)k * TS_YSPLIT + q) * n * KD + (r0 + rs) + (size_t)n * col];
    Bf[lr][col] = live ? y : mk(0.0, 0.0);
  }
  __syncthreads();
  z2* X = X_ + (size_t)k * n * KD;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int col = 8 * cq + c;
    z2 x = mk(0.0, 0.0);
#pragma unroll 8
    for (int i = 0; i < KD; ++i) x = x + Bf[lr][i] * Ts[i][col];
    if (!live) x = mk(0.0, 0.0);
    Xs[lr][col] = x;
    if (live) X[(r0 + r) + (size_t)n * col] = x;
  }
  __syncthreads();
  const z2* V = V_ + (size_t)k * n * KD;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int col = 8 * cq + c;
    Bf[lr][col] = ld_or_zero(V, (size_t)((r0 + r) + (size_t)n * col), live, (size_t)r0);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < KD * KD; e += 256) {
    const int i = e % KD, j = e / KD;
    z2 s = mk(0.0, 0.0);
#pragma unroll 8
    for (int l = 0; l < 64; ++l) s = s + cmul(Bf[l][i], Xs[l][j]);
    mpart_[(((size_t)k * nbxmax + b) * KD + j) * KD + i] = s;
  }
}

// Zm = T^H M, M = V^H X summed over the row blocks
__global__ void __launch_bounds__(256) k_ts_mz(i64 nbx, i64 nbxmax, const z2* mpart_, const z2* T_,
                                               z2* Zm_) {
  const int k = blockIdx.x;
  __shared__ z2 Ms[KD][KD + 1];
  __shared__ z2 Ts[KD][KD + 1];
  for (int e = threadIdx.x; e < KD * KD; e += 256) {
    const int i = e % KD, j = e / KD;
    z2 s = mk(0.0, 0.0);
    for (i64 b = 0; b < nbx; ++b) s = s + mpart_[(((size_t)k * nbxmax + b) * KD + j) * KD + i];
    Ms[i][j] = s;
    Ts[i][j] = T_[(size_t)k * KD * KD + e];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < KD * KD; e += 256) {
    const int i = e % KD, j = e / KD;
    z2 s = mk(0.0, 0.0);
    for (int l = 0; l < KD; ++l) s = s + cmul(Ts[l][i], Ms[l][j]);
    Zm_[(size_t)k * KD * KD + e] = s;
  }
}

// W = X - V Zm / 2 for 64 rows (W has the layout of the one-stage panel W)
__global__ void __launch_bounds__(256) k_ts_w(i64 n, i64 r0, i64 mp, const z2* X_, const z2* V_,
                                              const z2* Zm_, z2* W_) {
  const int k = blockIdx.y;
  const int lr = threadIdx.x / 4, cq = threadIdx.x % 4;
  const i64 r = 64 * (i64)blockIdx.x + lr;
  const bool live = r < mp;
  __shared__ z2 Zs[KD][KD + 1];
  __shared__ z2 Vr[64][KD + 1];
  for (int e = threadIdx.x; e < KD * KD; e += 256) Zs[e % KD][e / KD] = Zm_[(size_t)k * KD * KD + e];
  const z2* V = V_ + (size_t)k * n * KD;
  const z2* X = X_ + (size_t)k * n * KD;
  z2* W = W_ + (size_t)k * n * TRD_NB;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int col = 8 * cq + c;
    Vr[lr][col] = ld_or_zero(V, (size_t)((r0 + r) + (size_t)n * col), live, (size_t)r0);
  }
  __syncthreads();
  if (!live) return;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int col = 8 * cq + c;
    z2 s = mk(0.0, 0.0);
#pragma unroll 8
    for (int i = 0; i < KD; ++i) s = s + Vr[lr][i] * Zs[i][col];
    W[(r0 + r) + (size_t)n * col] = X[(r0 + r) + (size_t)n * col] - 0.5 * s;
  }
}

// ---------------------------------------------------------------- stage 2
// AB(c, d) = A(c + d, c) for d <= KD, zero for the bulge rows KD < d <= 2 KD
__global__ void k_ts_band(i64 n, const z2* A_, size_t sA, i64 lda, z2* AB_) {
  const int k = blockIdx.y;
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)n * LDAB) return;
  const i64 c = (i64)(e / LDAB), d = (i64)(e % LDAB);
  z2 a = mk(0.0, 0.0);
  if (d <= KD && c + d < n) {
    a = A_[k * sA + (c + d) + lda * c];
    if (d == 0) a.y = 0.0;
  }
  AB_[(size_t)k * n * LDAB + e] = a;
}

// One step (s, t = slot - 2 s) of sweep s per block (one wave), s = s_lo + blockIdx.x.
__global__ void __launch_bounds__(64) k_ts_chase(i64 n, i64 slot, i64 s_lo, i64 s_hi, z2* AB_,
                                                 z2* V2_, z2* tau2_, const i64* goff, i64 nblk2) {
  const i64 s = s_lo + blockIdx.x;
  if (s > s_hi) return;
  const int k = blockIdx.y, l = threadIdx.x;
  z2* AB = AB_ + (size_t)k * n * LDAB;
  const i64 t = slot - 2 * s;
  const i64 st = s + 1 + t * KD;
  const int L = (int)std::min<i64>(KD, n - st);  // rows st .. st + L - 1
  const i64 g = s / NBG;
  const int ii = (int)(s % NBG);
  const i64 j1 = st + L;                           // first row of the block below
  const bool below = j1 <= n - 1;                  // then L == KD
  const int lm = below ? (int)std::min<i64>(KD, n - j1) : 0;
  __shared__ z2 M[KD][KD + 1];  // the diagonal block, then the block below
  __shared__ z2 vs[KD], ws[KD], us[KD];
  __shared__ z2 tau_s;
  // prefetch the diagonal block (lower part, all lanes) and the block below
  z2 dv[16], bv[16];
#pragma unroll
  for (int h = 0; h < 16; ++h) {
    const int e = l + 64 * h, i = e % KD, j = e / KD;
    dv[h] = ld_or_zero(AB, (size_t)((st + j) * LDAB + (i - j)), i < L && j < L && i >= j,
                       (size_t)(st * LDAB));
    bv[h] = ld_or_zero(AB, (size_t)((st + j) * LDAB + (KD + i - j)), i < lm && j < L,
                       (size_t)(st * LDAB));
  }
  z2* V2 = V2_ + ((size_t)k * nblk2 + goff[g]) * NBG * KD;
  z2* tau2 = tau2_ + ((size_t)k * nblk2 + goff[g]) * NBG;
  if (t == 0) {
    // H_{s,0} from column s, rows st .. st + L - 1 (band offsets 1 .. L)
    const z2 x = ld_or_zero(AB, (size_t)(s * LDAB + 1 + l), l >= 1 && l < L, (size_t)(s * LDAB));
    const double xn2 = ts_wave_sum(abs2(x));
    const z2 alpha = AB[s * LDAB + 1];
    const TsRefl R = ts_reflector(alpha, xn2);
    if (l < KD) {
      const z2 v = l == 0 ? mk(1.0, 0.0) : (l < L ? x * R.scale : mk(0.0, 0.0));
      vs[l] = v;
      V2[(size_t)ii * KD + l] = v;
    }
    if (l == 0) {
      AB[s * LDAB + 1] = mk(R.beta, 0.0);
      tau2[ii] = R.tau;
      tau_s = R.tau;
    }
    if (l >= 1 && l < L) AB[s * LDAB + 1 + l] = mk(0.0, 0.0);
  } else {
    if (l < KD) vs[l] = V2[((size_t)t * NBG + ii) * KD + l];
    if (l == 0) tau_s = tau2[(size_t)t * NBG + ii];
  }
  // diagonal block D (Hermitian, both triangles in LDS)
#pragma unroll
  for (int h = 0; h < 16; ++h) {
    const int e = l + 64 * h, i = e % KD, j = e / KD;
    if (i >= j) {
      z2 a = dv[h];
      if (i == j) a.y = 0.0;
      M[i][j] = a;
      M[j][i] = conjz(a);
    }
  }
  __syncthreads();
  const z2 tau = tau_s;
  // The products below use both halves of the wave: lane l works on row (or
  // column) r = l % 32 over the half hh = l / 32 of the summation index, and
  // the two halves are added across lanes l and l + 32.
  const int r = l & (KD - 1), hh = l >> 5, q0 = 16 * hh;
  // D <- H^H D H:  y = D v, w = tau y - |tau|^2 (v^H y) v / 2, D -= w v^H + v w^H
  z2 y = mk(0.0, 0.0);
#pragma unroll
  for (int j = q0; j < q0 + 16; ++j) y = y + M[r][j] * vs[j];
  y = mk(y.x + __shfl_xor(y.x, 32, 64), y.y + __shfl_xor(y.y, 32, 64));
  const double sig = ts_wave_sum(hh == 0 ? (vs[r].x * y.x + vs[r].y * y.y) : 0.0);
  if (hh == 0) ws[r] = tau * y - (0.5 * abs2(tau) * sig) * vs[r];
  __syncthreads();
#pragma unroll
  for (int h = 0; h < 16; ++h) {
    const int e = l + 64 * h, i = e % KD, j = e / KD;
    if (i < L && j < L && i >= j) {
      z2 a = M[i][j] - mulc(ws[i], vs[j]) - mulc(vs[i], ws[j]);
      if (i == j) a.y = 0.0;
      AB[(st + j) * LDAB + (i - j)] = a;
    }
  }
  if (!below) return;
  __syncthreads();
  // the block below: B <- B H, then H_{s,t+1} from its first column, applied
  // from the left to its other columns
#pragma unroll
  for (int h = 0; h < 16; ++h) {
    const int e = l + 64 * h;
    M[e % KD][e / KD] = bv[h];
  }
  __syncthreads();
  {
    z2 u = mk(0.0, 0.0);
#pragma unroll
    for (int j = q0; j < q0 + 16; ++j) u = u + M[r][j] * vs[j];
    u = mk(u.x + __shfl_xor(u.x, 32, 64), u.y + __shfl_xor(u.y, 32, 64));
    if (hh == 0) us[r] = tau * u;
  }
  __syncthreads();
#pragma unroll
  for (int h = 0; h < 16; ++h) {
    const int e = l + 64 * h, i = e % KD, j = e / KD;
    M[i][j] = M[i][j] - mulc(us[i], vs[j]);
  }
  __syncthreads();
  const z2 x0 = (l >= 1 && l < lm) ? M[l][0] : mk(0.0, 0.0);
  const double xn2 = ts_wave_sum(abs2(x0));
  const TsRefl R = ts_reflector(M[0][0], xn2);
  const z2 v2 = l == 0 ? mk(1.0, 0.0) : ((l < lm) ? x0 * R.scale : mk(0.0, 0.0));
  __syncthreads();
  if (l < KD) {
    vs[l] = v2;  // H_{s,t+1}
    V2[((size_t)(t + 1) * NBG + ii) * KD + l] = v2;
  }
  if (l == 0) tau2[(size_t)(t + 1) * NBG + ii] = R.tau;
  __syncthreads();
  const z2 tb = conjz(R.tau);
  {
    // column r: q = v2^H B(:, r), B(:, r) -= conj(tau2) v2 q; column 0
    // becomes (beta, 0, ..., 0)
    z2 qv = mk(0.0, 0.0);
#pragma unroll
    for (int i = q0; i < q0 + 16; ++i) qv = qv + cmul(vs[i], M[i][r]);
    qv = mk(qv.x + __shfl_xor(qv.x, 32, 64), qv.y + __shfl_xor(qv.y, 32, 64));
    const z2 f = tb * qv;
    if (r == 0) {
#pragma unroll
      for (int i = q0; i < q0 + 16; ++i) M[i][0] = i == 0 ? mk(R.beta, 0.0) : mk(0.0, 0.0);
    } else if (r < L) {
#pragma unroll
      for (int i = q0; i < q0 + 16; ++i) M[i][r] = M[i][r] - vs[i] * f;
    }
  }
  __syncthreads();
#pragma unroll
  for (int h = 0; h < 16; ++h) {
    const int e = l + 64 * h, i = e % KD, j = e / KD;
    if (i < lm && j < L) AB[(st + j) * LDAB + (KD + i - j)] = M[i][j];
  }
}

__global__ void k_ts_de(i64 n, const z2* AB_, double* d_, double* e_) {
  const int k = blockIdx.y;
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const z2* AB = AB_ + (size_t)k * n * LDAB;
  d_[(size_t)k * n + i] = AB[i * LDAB].x;
  if (i < n - 1) e_[(size_t)k * n + i] = AB[i * LDAB + 1].x;
}

// Triangular factor of every block (g, t) of the chase's reflectors: in block
// rows, reflector i occupies rows i .. i + KD - 1; G(i, c) = v_i^H v_c over the
// shared rows, T(c, c) = tau_c, T(0:c, c) = -tau_c T(0:c, 0:c) G(0:c, c).
// Reflectors that the chase did not produce (past the end of their sweep)
// count as zero.
__global__ void __launch_bounds__(64) k_ts_t2(i64 n, i64 G, const i64* goff, const z2* V2_,
                                              const z2* tau2_, z2* T2_, i64 nblk2) {
  const int k = blockIdx.y, l = threadIdx.x;
  const i64 b = blockIdx.x;
  i64 lo = 0, hi = G - 1;  // the group of block b: goff[g] <= b < goff[g + 1]
  while (lo < hi) {
    const i64 mid = (lo + hi + 1) / 2;
    if (goff[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const i64 g = lo, t = b - goff[g];
  __shared__ z2 Vl[NBG][KD];
  __shared__ z2 Gs[NBG][NBG + 1];
  __shared__ z2 Ts[NBG][NBG + 1];
  __shared__ z2 tl[NBG];
  const z2* V2 = V2_ + ((size_t)k * nblk2 + b) * NBG * KD;
  for (int e = l; e < NBG * KD; e += 64) {
    const int i = e / KD;
    const i64 s = g * NBG + i;
    const bool ex = s <= n - 2 && t < ts_nsteps(n, s);
    Vl[i][e % KD] = ex ? V2[e] : mk(0.0, 0.0);
  }
  if (l < NBG) {
    const i64 s = g * NBG + l;
    const bool ex = s <= n - 2 && t < ts_nsteps(n, s);
    tl[l] = ex ? tau2_[((size_t)k * nblk2 + b) * NBG + l] : mk(0.0, 0.0);
  }
  __syncthreads();
  for (int e = l; e < NBG * NBG; e += 64) {
    const int i = e % NBG, c = e / NBG;
    z2 acc = mk(0.0, 0.0);
    if (i < c)
      for (int r = c; r < i + KD; ++r) acc = acc + cmul(Vl[i][r - i], Vl[c][r - c]);
    Gs[i][c] = acc;
  }
  __syncthreads();
  for (int c = 0; c < NBG; ++c) {
    if (l < c) {
      z2 acc = mk(0.0, 0.0);
      for (int q = l; q < c; ++q) acc = acc + Ts[l][q] * Gs[q][c];
      Ts[l][c] = (-1.0) * (tl[c] * acc);
    } else if (l == c) {
      Ts[c][c] = tl[c];
    } else if (l < NBG) {
      Ts[l][c] = mk(0.0, 0.0);
    }
    __syncthreads();
  }
  z2* T2 = T2_ + ((size_t)k * nblk2 + b) * NBG * NBG;
  for (int e = l; e < NBG * NBG; e += 64) T2[e] = Ts[e % NBG][e / NBG];
}

// ---------------------------------------------------------------- Q2 C
// Group g's blocks t = 0 .. cnt-1 applied to 64 columns of C (block), wave w
// owning 16 of them.  The 64-row window of C at base = g NB + 1 + t KD stays in
// the MFMA accumulators (4 row blocks of 16); a lane's value i of row block
// rb holds row 16 rb + row(lane, i), which is also the row that lane supplies
// as k-index lane / 16 of k-step i when the block is the B operand.  So
//   W  = V^H C    (reflectors x columns; V from LDS as the A operand)
//   W' = T W
//   C -= V W'
// run on the accumulators directly, skipping the 16 x 16 blocks where V or T
// is zero.  After each block the top 32 rows are final for the group and are
// stored; the window moves down by KD.  The next block's reflectors, its
// triangular factor and the 32 rows of C it adds are loaded into registers
// before the current block's products, so their latency overlaps them.  V's
// structural zeros (entries outside the parallelogram) are written to LDS
// once; each block overwrites only the KD x NB entries of its reflectors.
template <typename S, typename E>
__global__ void __launch_bounds__(256, 2) k_ts_q2(i64 n, i64 g, i64 cnt, i64 b0, i64 m, E* C_,
                                               size_t ldC, size_t sC, const z2* V2_, const z2* T2_,
                                               i64 nblk2) {
  typedef Mfma<S> MF;
  typedef typename MF::acc ACC;
  // FP64: Gauss's three-product complex form (a quarter fewer matrix-core
  // operations); FP32 keeps four products (its rounding is what bounds the
  // mixed path)
  typedef CAcc<S, sizeof(S) == 8> CA;
  const int k = blockIdx.y;
  E* C = C_ + k * sC;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, l16 = lane & 15;
  const i64 col = (i64)blockIdx.x * 64 + 16 * wv + l16;
  const bool cin = col < m;
  __shared__ S Vs[2][64][NBG + 1];
  __shared__ S Ts[2][NBG][NBG + 1];
  for (int e = tid; e < 64 * NBG; e += 256) {
    Vs[0][e % 64][e / 64] = (S)0;
    Vs[1][e % 64][e / 64] = (S)0;
  }
  ACC cre[4], cim[4];
  i64 base = g * NBG + 1;
  // prefetch registers: 4 reflector entries, 4 entries of T, 2 x 4 elements of C
  z2 pv[4], pt[4];
  bool pok[4];
  E pc[2][4];
  auto fetch_vt = [&](i64 tt) {
    const z2* V2 = V2_ + ((size_t)k * nblk2 + b0 + tt) * NBG * KD;
    const z2* T2 = T2_ + ((size_t)k * nblk2 + b0 + tt) * NBG * NBG;
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int e = tid + 256 * h, i = e / KD;  // reflector i, entry e % KD
      const i64 s = g * NBG + i;
      pok[h] = (s <= n - 2) & (tt < ts_nsteps(n, s));
      pv[h] = V2[pok[h] ? e : 0];
      pt[h] = T2[e];
    }
  };
  auto stage_vt = [&]() {
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int e = tid + 256 * h, i = e / KD, r = i + e % KD;
      const double f = pok[h] ? 1.0 : 0.0;
      Vs[0][r][i] = (S)(f * pv[h].x);
      Vs[1][r][i] = (S)(f * pv[h].y);
      Ts[0][e % NBG][e / NBG] = (S)pt[h].x;
      Ts[1][e % NBG][e / NBG] = (S)pt[h].y;
    }
  };
  // rows of row block rb of the window at bs (out-of-range elements read C[0]
  // and are zeroed when taken into the accumulators)
  auto fetch_rows = [&](E* dst, i64 bs, int rb) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const i64 r = bs + 16 * rb + MF::row(lane, i);
      dst[i] = C[(cin && r < n) ? r + ldC * col : 0];
    }
  };
  auto take_rows = [&](const E* src, i64 bs, int rb) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const i64 r = bs + 16 * rb + MF::row(lane, i);
      const S f = (cin && r < n) ? (S)1 : (S)0;
      cre[rb][i] = f * (S)src[i].x;
      cim[rb][i] = f * (S)src[i].y;
    }
  };
  auto store_rb = [&](int rb) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const i64 r = base + 16 * rb + MF::row(lane, i);
      if (cin && r < n) C[r + ldC * col] = narrow<E>(mk((double)cre[rb][i], (double)cim[rb][i]));
    }
  };
  fetch_vt(0);
  {
    E w0[4][4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) fetch_rows(w0[rb], base, rb);
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) take_rows(w0[rb], base, rb);
  }
  __syncthreads();  // the zeros of Vs are written
  stage_vt();
  __syncthreads();
  for (i64 t = 0; t < cnt; ++t, base += KD) {
    const bool more = t + 1 < cnt;
    if (more) {
      fetch_vt(t + 1);
      fetch_rows(pc[0], base + KD, 2);
      fetch_rows(pc[1], base + KD, 3);
    }
    // W = V^H C: reflector block mb meets row blocks mb .. mb + 2
    S wr[2][4], wi[2][4];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
      CA w;
      w.zero();
#pragma unroll
      for (int dr = 0; dr < 3; ++dr) {
        const int rb = mb + dr;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int rr = 16 * rb + MF::row(lane, i);
          w.mma(Vs[0][rr][16 * mb + l16], -Vs[1][rr][16 * mb + l16], cre[rb][i], cim[rb][i]);
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        wr[mb][i] = w.re(i);
        wi[mb][i] = w.im(i);
      }
    }
    // W' = T W (T upper triangular)
    S xr[2][4], xi[2][4];
#pragma unroll
    for (int mo = 0; mo < 2; ++mo) {
      CA x;
      x.zero();
#pragma unroll
      for (int mi = mo; mi < 2; ++mi)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int kk = 16 * mi + MF::row(lane, i);
          x.mma(Ts[0][16 * mo + l16][kk], Ts[1][16 * mo + l16][kk], wr[mi][i], wi[mi][i]);
        }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        xr[mo][i] = x.re(i);
        xi[mo][i] = x.im(i);
      }
    }
    // C -= V W': row block rb meets reflector blocks rb - 2 .. rb (within 0..1)
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      CA u;
      u.begin(cre[rb], cim[rb]);
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        if (rb < mb || rb > mb + 2) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int kk = 16 * mb + MF::row(lane, i);
          u.mma(-Vs[0][16 * rb + l16][kk], -Vs[1][16 * rb + l16][kk], xr[mb][i], xi[mb][i]);
        }
      }
      u.end(cre[rb], cim[rb]);
    }
    __syncthreads();  // every wave is done with this block's V and T
    store_rb(0);
    store_rb(1);
    if (more) {
      stage_vt();
      cre[0] = cre[2]; cim[0] = cim[2];
      cre[1] = cre[3]; cim[1] = cim[3];
      take_rows(pc[0], base + KD, 2);
      take_rows(pc[1], base + KD, 3);
      __syncthreads();
    } else {
      store_rb(2);
      store_rb(3);
    }
  }
}

// ---------------------------------------------------------------- host side
int twostage_reduce(TridiagWork& w, hipStream_t s, bool capturing) {
  TwoStageWork& ts = *w.ts;
  const i64 n = w.n;
  const int K = w.K;
  if (ts.n != n || ts.K < K) {
    set_error("twostage: workspace bound for another shape");
    return -2;
  }
  const i64 nbqmax = nbq_of(n), nbxmax = nbx_of(n);
  long launches = 0;
  auto throttle = [&]() -> int {
    // bound the launches in flight when not captured (see enqueue_reduction)
    if (!capturing && (++launches % 1024) == 0) FORGE_HIP(hipStreamSynchronize(s));
    return 0;
  };
  for (i64 i0 = 0;; i0 += KD) {
    const i64 r0 = i0 + KD, mp = n - r0;
    if (mp <= 0) break;
    const int nq = (int)std::min<i64>(KD, mp);
    const unsigned nbq = (unsigned)nbq_of(mp);
    hipLaunchKernelGGL(k_ts_qnorm, dim3(nbq, K), dim3(256), 0, s, n, i0, w.A, w.sA, w.lda, ts.npart,
                       nbqmax);
    FORGE_LAUNCH();
    for (int jj = 0; jj < nq; ++jj) {
      hipLaunchKernelGGL(k_ts_qa, dim3(nbq, K), dim3(256), 0, s, n, i0, jj, nq, w.A, w.sA, w.lda,
                         ts.V, w.tau, ts.npart, ts.qpart, ts.gpart, nbqmax);
      FORGE_LAUNCH();
      hipLaunchKernelGGL(k_ts_qb, dim3(nbq, K), dim3(256), 0, s, n, i0, jj, nq, w.A, w.sA, w.lda,
                         ts.V, w.tau, ts.qpart, ts.npart, nbqmax);
      FORGE_LAUNCH();
      if (throttle()) return -1;
    }
    hipLaunchKernelGGL(k_ts_tbuild, dim3(K), dim3(256), 0, s, n, i0, nq, w.tau, ts.gpart,
                       (i64)nbq, nbqmax, ts.T);
    FORGE_LAUNCH();
    // column splits of A22 V: enough blocks to fill the chip several times
    const i64 ntile = (mp + 63) / 64;
    i64 nsplit = std::max<i64>(1, std::min<i64>(TS_YSPLIT, 2048 / std::max<i64>(1, ntile * K)));
    i64 KS = ((mp + nsplit - 1) / nsplit + HS - 1) / HS * HS;
    nsplit = (mp + KS - 1) / KS;
    hipLaunchKernelGGL(k_ts_hemm, dim3((unsigned)(ntile * nsplit), K), dim3(256), 0, s, n, r0, mp,
                       nq, w.A, w.sA, w.lda, ts.V, ts.Y, ntile, KS);
    FORGE_LAUNCH();
    const unsigned nbx = (unsigned)nbx_of(mp);
    hipLaunchKernelGGL(k_ts_x, dim3(nbx, K), dim3(256), 0, s, n, r0, mp, (int)nsplit, ts.Y, ts.T,
                       ts.V, ts.X, ts.mpart, nbxmax);
    FORGE_LAUNCH();
    hipLaunchKernelGGL(k_ts_mz, dim3(K), dim3(256), 0, s, (i64)nbx, nbxmax, ts.mpart, ts.T, ts.Zm);
    FORGE_LAUNCH();
    hipLaunchKernelGGL(k_ts_w, dim3(nbx, K), dim3(256), 0, s, n, r0, mp, ts.X, ts.V, ts.Zm, w.W);
    FORGE_LAUNCH();
    {
      const int r = her2k_twostage(w, r0, i0, nq, s);
      if (r) return r;
    }
    if (throttle()) return -1;
  }
  hipLaunchKernelGGL(k_ts_band, dim3(cdiv(n * LDAB, 256), K), dim3(256), 0, s, n, w.A, w.sA, w.lda,
                     ts.AB);
  FORGE_LAUNCH();
  for (i64 slot = 0; slot <= 2 * (n - 2); ++slot) {
    const i64 num = KD * slot - n + 2, den = 2 * KD - 1;
    const i64 s_lo = num > 0 ? (num + den - 1) / den : 0;
    const i64 s_hi = std::min<i64>(slot / 2, n - 2);
    if (s_lo > s_hi) continue;
    hipLaunchKernelGGL(k_ts_chase, dim3((unsigned)(s_hi - s_lo + 1), K), dim3(64), 0, s, n, slot,
                       s_lo, s_hi, ts.AB, ts.V2, ts.tau2, ts.goff, ts.nblk2);
    FORGE_LAUNCH();
    if (throttle()) return -1;
  }
  hipLaunchKernelGGL(k_ts_de, dim3(cdiv(n, 256), K), dim3(256), 0, s, n, ts.AB, w.d, w.e);
  FORGE_LAUNCH();
  if (ts.nblk2 > 0) {
    hipLaunchKernelGGL(k_ts_t2, dim3((unsigned)ts.nblk2, K), dim3(64), 0, s, n, ts.G, ts.goff,
                       ts.V2, ts.tau2, ts.T2, ts.nblk2);
    FORGE_LAUNCH();
  }
  w.two_stage = 1;
  w.voff = KD;
  return 0;
}

template <typename S, typename E>
static int apply_q2_t(const TridiagWork& t, i64 m, E* C, size_t ldC, size_t sC, hipStream_t s) {
  const TwoStageWork& ts = *t.ts;
  if (m <= 0 || t.n < 2) return 0;
  const unsigned nbc = cdiv(m, 64);
  const i64 n = t.n;
  for (i64 g = ts.G - 1; g >= 0; --g) {
    const i64 cnt = ts.hoff[g + 1] - ts.hoff[g];
    if (ts.q2_timed && g % TS_Q2_STRIDE == 0) {
      const size_t q = ts.q2rec++;
      while (ts.q2ev.size() < 2 * (q + 1)) {
        hipEvent_t e;
        FORGE_HIP(hipEventCreate(&e));
        ts.q2ev.push_back(e);
      }
      // reflector elements of the group: sweep s's reflectors cover rows s + 1 .. n - 1
      double elems = 0.0;
      for (i64 sw = g * NBG; sw < std::min<i64>(g * NBG + NBG, n - 1); ++sw) elems += (double)(n - 1 - sw);
      if (ts.q2flops.size() <= q) ts.q2flops.resize(q + 1);
      if (ts.q2bytes.size() <= q) ts.q2bytes.resize(q + 1);
      ts.q2flops[q] = 16.0 * elems * (double)m * t.K;
      // the eigenvector rows the group's steps pass over, read and written once
      ts.q2bytes[q] = 2.0 * (double)(n - 1 - g * NBG) * (double)m * t.K * sizeof(E);
      hipExtLaunchKernelGGL((k_ts_q2<S, E>), dim3(nbc, t.K), dim3(256), 0, s, ts.q2ev[2 * q],
                            ts.q2ev[2 * q + 1], 0, n, g, cnt, ts.hoff[g], m, C, ldC, sC, ts.V2, ts.T2,
                            ts.nblk2);
    } else {
      hipLaunchKernelGGL((k_ts_q2<S, E>), dim3(nbc, t.K), dim3(256), 0, s, n, g, cnt, ts.hoff[g], m,
                         C, ldC, sC, ts.V2, ts.T2, ts.nblk2);
    }
    FORGE_LAUNCH();
  }
  return 0;
}

int twostage_q2_timing(const TwoStageWork& ts, long* launches, double* ms, double* flops,
                       double* bytes) {
  *launches = 0;
  *ms = 0.0;
  *flops = 0.0;
  *bytes = 0.0;
  if (ts.q2rec == 0) return 0;
  FORGE_HIP(hipEventSynchronize(ts.q2ev[2 * ts.q2rec - 1]));
  for (size_t q = 0; q < ts.q2rec; ++q) {
    float a = 0.f;
    FORGE_HIP(hipEventElapsedTime(&a, ts.q2ev[2 * q], ts.q2ev[2 * q + 1]));
    *ms += a;
    *flops += ts.q2flops[q];
    *bytes += ts.q2bytes[q];
  }
  *launches = (long)ts.q2rec;
  ts.q2rec = 0;
  return 0;
}

void twostage_release_events(TwoStageWork& ts) {
  for (hipEvent_t e : ts.q2ev) (void)hipEventDestroy(e);
  ts.q2ev.clear();
  ts.q2flops.clear();
  ts.q2bytes.clear();
  ts.q2rec = 0;
}

int twostage_apply_q2(const TridiagWork& t, i64 m, z2* C, size_t ldC, size_t sC, hipStream_t s) {
  return apply_q2_t<double, z2>(t, m, C, ldC, sC, s);
}
int twostage_apply_q2(const TridiagWork& t, i64 m, float2* C, size_t ldC, size_t sC,
                      hipStream_t s) {
  return apply_q2_t<float, float2>(t, m, C, ldC, sC, s);
}

}  // namespace forge
