// Batched reduction of K Hermitian matrices (lower triangle, column-major,
// leading dimension lda) to real symmetric tridiagonal form A = Q T Q^H,
// Q = H_0 H_1 ... H_{n-2}, H_j = I - tau_j v_j v_j^H.  This is the blocked
// lower Householder reduction of LAPACK zhetrd/zlatrd (the reference reaches
// it through zheevr, LapackRoutines.f90:414-426) with the same per-column
// arithmetic:
//
//   column update   a(j:n) -= A(j:n, i0:j) conj(W(j, :)) + W(j:n, :) conj(A(j, i0:j))
//   reflector       beta = -sign(Re alpha) |(alpha, x)|, tau = (beta - alpha)/beta,
//                   v = (1, x / (alpha - beta)), e_j = beta
//   w               = tau (A22 v - V (W^H v) - W (V^H v)),   w += -tau/2 (w^H v) v
//   panel end       A22 -= V W^H + W V^H   (lower triangle, k_trd_her2k_mfma)
//
// Each column is three launches over all K matrices, with no atomics and no
// device-scope fences (the L2 caches of the XCDs are not coherent with each
// other, so every such fence would write back L2): a producer kernel leaves
// per-block partial sums and every block of the consumer kernel reduces the
// few partials it needs itself, in a fixed order.
//   k_trd_col    correction of the previous w (-tau/2 (w^H v) v, own rows), the
//                column update, norm partials of the new column
//   k_trd_hemv   reflector scalars from the norm partials; A22 v with the lower
//                triangle of A22 read once (32-column strips); W^H v, V^H v
//                partials in dedicated blocks
//   k_trd_fin    w = tau (A22 v - V W^H v - W V^H v), v written back, w^H v
//                partials
// The correction of the panel's last w is applied on the fly by k_trd_her2k_mfma.
//
// Mixed precision (TridiagWork::S set): the hemv streams an FP32 copy of the
// trailing matrix, written by k_trd_to_fp32 before the first panel and by
// k_trd_her2k_mfma after every trailing update, and accumulates in FP64 against the
// FP64 Householder vector.  The hemv is bound by HBM bandwidth, so halving its
// bytes halves its time; the reduction's backward error becomes that of the
// FP32 rounding of A22 (|dA| <= 2^-24 |A|), while every other quantity (the
// panel, W, tau, d, e) stays FP64.
#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "kernels.hpp"

namespace forge {

constexpr int NB = TRD_NB;
constexpr int TS = TRD_TS;
constexpr int SEGR = TRD_SEGR;  // the smaller segment height: sizes the partial arrays

// column strips of width TS over the trailing order m
__host__ __device__ static inline i64 nt_of(i64 m) { return (m + TS - 1) / TS; }
// row segments of strip J (rows 32J .. m-1) of height segr
__host__ __device__ static inline i64 nseg_of(i64 m, i64 J, i64 segr = SEGR) {
  return (m - TS * J + segr - 1) / segr;
}
__host__ __device__ static inline i64 nblk_of(i64 m) { return (m + 255) / 256; }
// rows per k_trd_fin block and its block count
constexpr int FIN_ROWS = 64;
__host__ __device__ static inline i64 nfin_of(i64 m) { return (m + FIN_ROWS - 1) / FIN_ROWS; }
// per-matrix stride of the norm (k_trd_col) and dot (k_trd_fin) partials
__host__ __device__ static inline i64 npart_stride(i64 n) { return nfin_of(n) + 1; }

struct TrdLayout {
  size_t A_d, A_e, A_tau, A_W, A_P, A_P2, A_tp, A_np, A_dp, A_sc, total;
};

static TrdLayout layout(i64 n, int K) {
  i64 nt = nt_of(n), ms = nseg_of(n, 0), nb = npart_stride(n);
  TrdLayout L;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  L.A_d = take(sizeof(double) * K * n);
  L.A_e = take(sizeof(double) * K * n);
  L.A_tau = take(sizeof(z2) * K * n);
  L.A_W = take(sizeof(z2) * K * n * NB);
  L.A_P = take(sizeof(z2) * (size_t)K * nt * n);
  L.A_P2 = take(sizeof(z2) * (size_t)K * nt * ms * TS);
  L.A_tp = take(sizeof(z2) * (size_t)K * ms * TRD_PB * 2 * NB);
  L.A_np = take(sizeof(double) * (size_t)K * nb);
  L.A_dp = take(sizeof(z2) * (size_t)K * nb);
  L.A_sc = take(sizeof(z2) * (size_t)K * 4);
  L.total = off;
  return L;
}

size_t tridiag_workspace_bytes(i64 n, int K) { return layout(n, K).total; }

int tridiag_bind(TridiagWork& w, void* base, i64 n, int K) {
  TrdLayout L = layout(n, K);
  char* b = (char*)base;
  w.n = n;
  w.K = K;
  w.lda = n;
  w.d = (double*)(b + L.A_d);
  w.e = (double*)(b + L.A_e);
  w.tau = (z2*)(b + L.A_tau);
  w.W = (z2*)(b + L.A_W);
  w.P = (z2*)(b + L.A_P);
  w.P2 = (z2*)(b + L.A_P2);
  w.tpart = (z2*)(b + L.A_tp);
  w.npart = (double*)(b + L.A_np);
  w.dpart = (z2*)(b + L.A_dp);
  w.scal = (z2*)(b + L.A_sc);
  w.S = nullptr;
  w.sS = 0;
  w.evs = nullptr;
  w.nev = 0;
  w.first_col = nullptr;
  w.use_graph = 0;
  w.graph = nullptr;
  w.ts = nullptr;
  w.voff = 1;
  w.two_stage = 0;
  return 0;
}

// ---------------------------------------------------------------- helpers
__device__ inline double block_sum256(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

struct Reflector {
  z2 tau, scale;
  double beta;
};

// zlarfg of the column j: alpha = A(j+1, j), |x|^2 = xn2 (rows j+2..).
__device__ inline Reflector reflector(z2 alpha, double xn2) {
  Reflector R;
  if (xn2 == 0.0 && alpha.y == 0.0) {
    R.tau = mk(0.0, 0.0);
    R.beta = alpha.x;
    R.scale = mk(0.0, 0.0);
  } else {
    double nr = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
    R.beta = alpha.x >= 0.0 ? -nr : nr;
    R.tau = mk((R.beta - alpha.x) / R.beta, -alpha.y / R.beta);
    z2 den = mk(alpha.x - R.beta, alpha.y);  // 1 / (alpha - beta)
    double dd = abs2(den);
    R.scale = mk(den.x / dd, -den.y / dd);
  }
  return R;
}

// Sum of nb partials by the first wave: lane l adds partials l, l + 64, ...,
// then a fixed butterfly; one memory latency instead of nb dependent loads.
__device__ inline double wave_sum_partials(const double* part, unsigned nb) {
  double s = 0.0;
  for (unsigned b = threadIdx.x; b < nb; b += 64) s += part[b];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  return s;
}

// Reflector of column j from the norm partials of k_trd_col (fixed order).
// All threads of the block return the same value.
__device__ inline Reflector reflector_of(const z2* A, i64 lda, i64 n, i64 j, const double* npart) {
  __shared__ Reflector Rs;
  if (threadIdx.x < 64) {
    const double xn2 = wave_sum_partials(npart, (unsigned)nblk_of(n - j));
    if (threadIdx.x == 0) Rs = reflector(A[(j + 1) + lda * j], xn2);
  }
  __syncthreads();
  Reflector R = Rs;
  __syncthreads();
  return R;
}

// -tau_j/2 (w_j^H v_j) of the finished column j from the partials of
// k_trd_fin (its grid had nfin_of(n - j - 1) blocks).
__device__ inline z2 alpha_of(i64 n, i64 j, z2 tau, const z2* dpart) {
  __shared__ z2 as;
  if (threadIdx.x < 64) {
    const unsigned nb = (unsigned)nfin_of(n - j - 1);
    double sx = 0.0, sy = 0.0;
    for (unsigned b = threadIdx.x; b < nb; b += 64) {
      sx += dpart[b].x;
      sy += dpart[b].y;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      sx += __shfl_xor(sx, off, 64);
      sy += __shfl_xor(sy, off, 64);
    }
    if (threadIdx.x == 0) as = (-0.5) * (tau * mk(sx, sy));
  }
  __syncthreads();
  z2 a = as;
  __syncthreads();
  return a;
}

// ---------------------------------------------------------------- K1: column update
__global__ void __launch_bounds__(256) k_trd_col(i64 n, i64 j, i64 i0, int jj, z2* A_, size_t sA,
                                                 i64 lda, z2* W_, double* d_, const z2* tau_,
                                                 double* npart_, const z2* dpart_, i64 nbmax) {
  const int k = blockIdx.y;
  z2* A = A_ + k * sA;
  z2* W = W_ + (size_t)k * n * NB;
  __shared__ z2 arow[NB], wrow[NB];
  __shared__ double sh[256];
  // pending correction of column jj-1 (reflector j-1): w += alpha v
  z2 alpha = mk(0.0, 0.0);
  if (jj > 0) alpha = alpha_of(n, j - 1, tau_[(size_t)k * n + j - 1], dpart_ + (size_t)k * nbmax);
  if (threadIdx.x < jj) {
    const int q = threadIdx.x;
    arow[q] = A[j + lda * (i0 + q)];
    z2 wq = W[j + n * q];
    if (q == jj - 1) wq = wq + alpha;  // v_{j-1} has its unit element in row j
    wrow[q] = wq;
  }
  __syncthreads();
  const i64 r = j + (i64)blockIdx.x * 256 + threadIdx.x;
  double nrm = 0.0;
  if (r < n) {
    z2 a = A[r + lda * j];
    if (r == j) a.y = 0.0;
    if (jj > 0) {
      z2 wc = W[r + n * (jj - 1)];
      wc = wc + alpha * A[r + lda * (j - 1)];
      if (r > j) W[r + n * (jj - 1)] = wc;  // row j is not read again
      for (int q = 0; q < jj - 1; ++q) {
        a = a - mulc(A[r + lda * (i0 + q)], wrow[q]);
        a = a - mulc(W[r + n * q], arow[q]);
      }
      a = a - mulc(A[r + lda * (j - 1)], wrow[jj - 1]);
      a = a - mulc(wc, arow[jj - 1]);
    }
    if (r == j) {
      a.y = 0.0;
      d_[(size_t)k * n + j] = a.x;
    }
    A[r + lda * j] = a;
    if (r >= j + 2) nrm = abs2(a);
  }
  nrm = block_sum256(nrm, sh);
  if (threadIdx.x == 0) npart_[(size_t)k * nbmax + blockIdx.x] = nrm;
}

// ---------------------------------------------------------------- K2: hemv
// grid (PB + nt, nseg(m, 0, SG), K).  Strip blocks (x = PB + J) stream column
// strip J (32 columns, c0 = base + 32 J) over rows
// [32 J + s SG, min(m, 32 J + (s+1) SG)) of the trailing matrix:
//   P[J][r]      = sum_c A(r, c) v_c                (c <= r inside the diagonal tile)
//   P2[J][s][c]  = sum_r conj(A(r, c)) v_r          (r >  c inside the diagonal tile)
// Blocks x < PB form the W^H v and V^H v partials of chunk x of row segment s.
// TM: element type of the streamed matrix (z2: A itself; float2: the FP32
// copy).  SG: rows per segment, chosen per trailing order by the tuner below
// (k_trd_fin sums the partials of the same segmentation).

// W^H v and V^H v over rows [rs, re) of the trailing part: chunk pb of
// segment s.  These blocks come first in the grid and are small, so they
// finish under the strip blocks instead of forming the tail.
template <int SG>
__device__ __forceinline__ void hemv_panel_block(i64 n, i64 j, i64 i0, int jj, const z2* A, i64 lda,
                                                 const z2* W_, const double* npart, z2* scal_,
                                                 z2* tpart_, i64 msmax, int k, i64 s) {
  const i64 base = j + 1, m = n - base;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int pb = blockIdx.x;
  const i64 rs = s * SG + pb * (SG / TRD_PB);
  const i64 re = rs + SG / TRD_PB < m ? rs + SG / TRD_PB : m;
  const bool publish = s == 0 && pb == 0;
  Reflector R{};
  if (publish || (jj > 0 && rs < m))  // uniform over the block
    R = reflector_of(A, lda, n, j, npart);
  if (publish && threadIdx.x == 0) {  // published for k_trd_fin, which rewrites A(j+1, j)
    scal_[k * 4 + 0] = R.tau;
    scal_[k * 4 + 1] = R.scale;
    scal_[k * 4 + 2] = mk(R.beta, 0.0);
  }
  const z2* W = W_ + (size_t)k * n * NB;
  constexpr int QW = NB / 4;  // panel columns per wave, lanes over rows
  const int qb = wv * QW;
  // two panel columns per pass keep this branch's registers below the
  // strip branch's, which sets the kernel's occupancy
  z2* tp = tpart_ + (((size_t)k * msmax + s) * TRD_PB + pb) * 2 * NB;
  for (int q0 = 0; q0 < QW; q0 += 2) {
    z2 t1[2] = {mk(0.0, 0.0), mk(0.0, 0.0)}, t2[2] = {mk(0.0, 0.0), mk(0.0, 0.0)};
    if (qb + q0 < jj) {
      for (i64 rr = rs + lane; rr < re; rr += 64) {
        const i64 r = base + rr;
        const z2 vr = (rr == 0) ? mk(1.0, 0.0) : A[r + lda * j] * R.scale;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          if (qb + q0 + q < jj) {
            t1[q] = t1[q] + cmul(W[r + n * (qb + q0 + q)], vr);
            t2[q] = t2[q] + cmul(A[r + lda * (i0 + qb + q0 + q)], vr);
          }
        }
      }
    }
    // written also for empty chunks: k_trd_fin sums every chunk of the
    // segments that hold rows
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      double a = t1[q].x, b = t1[q].y, c = t2[q].x, d = t2[q].y;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c += __shfl_xor(c, off, 64);
        d += __shfl_xor(d, off, 64);
      }
      if (lane == 0) {
        tp[qb + q0 + q] = mk(a, b);
        tp[NB + qb + q0 + q] = mk(c, d);
      }
    }
  }
}

// v of strip J's columns and of the block's rows (relative rows ra0 + q;
// zero outside [rs, re)), staged once in LDS
template <int SG>
__device__ __forceinline__ void hemv_stage_v(const z2* A, i64 lda, i64 n, i64 j, i64 base, i64 c0,
                                             i64 rs, i64 re, i64 ra0, z2 scale, z2* vJ, z2* vrow) {
  if (threadIdx.x < TS) {
    const i64 c = c0 + threadIdx.x;
    vJ[threadIdx.x] = c >= n ? mk(0.0, 0.0) : (c == base ? mk(1.0, 0.0) : A[c + lda * j] * scale);
  }
  for (int q = threadIdx.x; q < SG + 32; q += 256) {
    const i64 rr = ra0 + q;
    vrow[q] = (rr < rs || rr >= re) ? mk(0.0, 0.0)
                                    : (rr == 0 ? mk(1.0, 0.0) : A[(base + rr) + lda * j] * scale);
  }
}

// Layout 0 (FP64 matrix or FP32 copy).  Lane l of a wave covers row
// (l mod 16) and the 8 columns 8 (l / 16) .. + 7 of a 16-row step, so one
// load instruction reads four column segments of 16 rows (256 bytes in
// FP64, 128 in FP32), a thread holds only 8 column partials, and the 4 waves
// take interleaved 16-row steps.  The FP32 copy is read one step ahead
// (8-byte loads alone leave too few bytes in flight per wave); the FP64
// loads gain nothing from it.
template <typename TM, int SG>
__device__ __forceinline__ void hemv_strip16(i64 n, i64 j, const z2* A, i64 lda, const TM* M,
                                             const double* npart, z2* P_, z2* P2_, i64 ntmax,
                                             i64 msmax, int k, i64 s) {
  const i64 base = j + 1, m = n - base, nt = nt_of(m);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const i64 J = (i64)blockIdx.x - TRD_PB;  // column strip
  const i64 rs = TS * J + s * SG;  // first row (relative to base) of this block
  if (J >= nt || rs >= m) return;
  const i64 re = rs + SG < m ? rs + SG : m;
  const Reflector R = reflector_of(A, lda, n, j, npart);
  const i64 c0 = base + J * TS;
  const int ncol = (int)((n - c0) < TS ? (n - c0) : TS);
  __shared__ z2 vJ[TS];
  __shared__ z2 red[4][TS];
  __shared__ z2 vrow[SG + 32];
  const i64 ra0 = ((base + rs) & ~(i64)15) - base;  // may be < rs (masked rows)
  hemv_stage_v<SG>(A, lda, n, j, base, c0, rs, re, ra0, R.scale, vJ, vrow);
  __syncthreads();
  const int rsub = lane & 15, cg = lane >> 4;  // row in the step, column group
  const int cb = cg * 8;                       // first column of this lane's group
  z2 acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = mk(0.0, 0.0);
  z2* P = P_ + ((size_t)k * ntmax + J) * n;
  // uniform base (strip start) and 32-bit per-lane offsets: the loads use a
  // scalar base, which saves a VGPR pair per column address; with lda a
  // multiple of 16 every 16-row column chunk is one aligned 256-byte segment
  const TM* strip = M + (size_t)lda * c0;
  const unsigned lda32 = (unsigned)lda, cofs = (unsigned)cb * lda32;
  constexpr bool PF = sizeof(TM) == 8;
  TM nxt[8];
  if (PF) {
    const i64 rr = ra0 + 16 * wv + rsub;
    const bool ok = rr >= rs && rr < re;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      nxt[q] = (ok && cb + q < ncol) ? strip[cofs + (unsigned)(base + rr) + lda32 * q] : TM{};
  }
  for (i64 rr0 = ra0 + 16 * wv; rr0 < re; rr0 += 64) {
    const i64 rr = rr0 + rsub;  // row relative to base
    const bool valid = rr >= rs && rr < re;
    const i64 r = base + rr;
    const z2 vr = vrow[rr - ra0];  // zero unless valid
    const i64 dr = rr - TS * J;  // < 32 inside the diagonal tile
    z2 a[8];
    if (PF) {
#pragma unroll
      for (int q = 0; q < 8; ++q) a[q] = widen(nxt[q]);
      const i64 rn = rr + 64;
      const bool ok = rn >= rs && rn < re;
#pragma unroll
      for (int q = 0; q < 8; ++q)
        nxt[q] = (ok && cb + q < ncol) ? strip[cofs + (unsigned)(base + rn) + lda32 * q] : TM{};
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q)
        a[q] = (valid && cb + q < ncol) ? widen(strip[cofs + (unsigned)r + lda32 * q]) : mk(0.0, 0.0);
    }
    z2 y = mk(0.0, 0.0);
    if (rr0 >= TS * J + TS) {
      // below the diagonal tile (wave-uniform): every element counts in both
      // products (rows outside [rs, re) were loaded as zeros)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        y = y + a[q] * vJ[cb + q];
        acc[q] = acc[q] + cmul(a[q], vr);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const z2 zero = mk(0.0, 0.0);
        const int c = cb + q;
        y = y + ((c <= dr) ? a[q] : zero) * vJ[c];
        acc[q] = acc[q] + cmul((c < dr) ? a[q] : zero, vr);
      }
    }
    // row sum over the 4 column groups (lanes l, l^16, l^32, l^48)
    y.x += __shfl_xor(y.x, 16, 64);
    y.y += __shfl_xor(y.y, 16, 64);
    y.x += __shfl_xor(y.x, 32, 64);
    y.y += __shfl_xor(y.y, 32, 64);
    if (cg == 0 && valid) P[rr] = y;
  }
  // column partials: sum over the 16 row lanes, then over the 4 waves
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    double x = acc[q].x, yv = acc[q].y;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      x += __shfl_xor(x, off, 64);
      yv += __shfl_xor(yv, off, 64);
    }
    if (rsub == 0) red[wv][cb + q] = mk(x, yv);
  }
  __syncthreads();
  if (threadIdx.x < TS) {
    z2 t = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
    P2_[(((size_t)k * ntmax + J) * msmax + s) * TS + threadIdx.x] = t;
  }
}

// 16-byte load of rows rr, rr + 1 (relative to base) of one column of the
// FP32 copy; rows outside [rs, re) and columns past the strip read as zero
__device__ __forceinline__ float4 load_pair32(const float2* p, bool colok, bool ok0, bool ok1) {
  if (colok && ok0 && ok1) return *(const float4*)p;
  const float2 a = (colok && ok0) ? p[0] : make_float2(0.f, 0.f);
  const float2 b = (colok && ok1) ? p[1] : make_float2(0.f, 0.f);
  return make_float4(a.x, a.y, b.x, b.y);
}

// Layout 1 (FP32 copy).  Lane l of a wave covers the rows 2 (l mod 8), + 1
// of a 16-row step and the 4 columns 4 (l / 8) .. + 3 of the strip: each load is 16
// bytes (two rows of one column), 8 lanes read one 128-byte column segment,
// and one load instruction reads 8 of them.  A lane holds 4 column partials,
// and the loads of the next two steps are in flight while a step is summed.  Accumulation is FP64 against the FP64 v, as for the FP64
// matrix.
template <int SG>
__device__ __forceinline__ void hemv_strip8(i64 n, i64 j, const z2* A, i64 lda, const float2* M,
                                             const double* npart, z2* P_, z2* P2_, i64 ntmax,
                                             i64 msmax, int k, i64 s) {
  const i64 base = j + 1, m = n - base, nt = nt_of(m);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const i64 J = (i64)blockIdx.x - TRD_PB;  // column strip
  const i64 rs = TS * J + s * SG;  // first row (relative to base) of this block
  if (J >= nt || rs >= m) return;
  const i64 re = rs + SG < m ? rs + SG : m;
  const Reflector R = reflector_of(A, lda, n, j, npart);
  const i64 c0 = base + J * TS;
  const int ncol = (int)((n - c0) < TS ? (n - c0) : TS);
  __shared__ z2 vJ[TS];
  __shared__ z2 red[4][TS];
  __shared__ z2 vrow[SG + 32];
  // row steps start at an absolute multiple of 16 rows, so every row pair
  // starts at an even row and (lda even) every 16-byte load is aligned
  const i64 ra0 = ((base + rs) & ~(i64)15) - base;  // may be < rs (masked rows)
  hemv_stage_v<SG>(A, lda, n, j, base, c0, rs, re, ra0, R.scale, vJ, vrow);
  __syncthreads();
  const int rp = lane & 7, cg = lane >> 3;  // row pair in the step, column group
  const int cb = cg * 4;                    // first column of this lane's group
  // the strip columns' v stays in LDS (a broadcast read per use)
  const z2* vc = vJ + cb;
  z2 acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = mk(0.0, 0.0);
  z2* P = P_ + ((size_t)k * ntmax + J) * n;
  const float2* strip = M + (size_t)lda * c0;
  const unsigned lda32 = (unsigned)lda, cofs = (unsigned)cb * lda32;
  // loads of the lane's row pair at relative row rn into d
  auto load_step = [&](i64 rn, float4* d) {
    const bool ok0 = rn >= rs && rn < re, ok1 = rn + 1 >= rs && rn + 1 < re;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      d[q] = load_pair32(strip + cofs + (unsigned)(base + rn) + lda32 * q, cb + q < ncol, ok0, ok1);
  };
  // two steps in flight ahead of the one being summed
  float4 pf0[4], pf1[4];
  load_step(ra0 + 16 * wv + 2 * rp, pf0);
  load_step(ra0 + 16 * wv + 2 * rp + 64, pf1);
  for (i64 rr0 = ra0 + 16 * wv; rr0 < re; rr0 += 64) {
    const i64 rr = rr0 + 2 * rp;  // first of the lane's two rows (relative to base)
    float4 cur[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      cur[q] = pf0[q];
      pf0[q] = pf1[q];
    }
    load_step(rr + 128, pf1);
    const z2 vr0 = vrow[rr - ra0], vr1 = vrow[rr + 1 - ra0];  // zero unless valid
    const i64 dr = rr - TS * J;  // < 32 inside the diagonal tile
    z2 y0 = mk(0.0, 0.0), y1 = mk(0.0, 0.0);
    if (rr0 >= TS * J + TS) {
      // below the diagonal tile (wave-uniform): every element counts
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const z2 a0 = mk((double)cur[q].x, (double)cur[q].y);
        const z2 a1 = mk((double)cur[q].z, (double)cur[q].w);
        y0 = y0 + a0 * vc[q];
        y1 = y1 + a1 * vc[q];
        acc[q] = acc[q] + cmul(a0, vr0) + cmul(a1, vr1);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const z2 zero = mk(0.0, 0.0);
        const int c = cb + q;
        const z2 a0 = mk((double)cur[q].x, (double)cur[q].y);
        const z2 a1 = mk((double)cur[q].z, (double)cur[q].w);
        y0 = y0 + ((c <= dr) ? a0 : zero) * vc[q];
        y1 = y1 + ((c <= dr + 1) ? a1 : zero) * vc[q];
        acc[q] = acc[q] + cmul((c < dr) ? a0 : zero, vr0) + cmul((c < dr + 1) ? a1 : zero, vr1);
      }
    }
    // row sums over the 8 column groups (lanes l, l^8, ..., l^56)
#pragma unroll
    for (int off = 8; off < 64; off <<= 1) {
      y0.x += __shfl_xor(y0.x, off, 64);
      y0.y += __shfl_xor(y0.y, off, 64);
      y1.x += __shfl_xor(y1.x, off, 64);
      y1.y += __shfl_xor(y1.y, off, 64);
    }
    if (cg == 0) {
      if (rr >= rs && rr < re) P[rr] = y0;
      if (rr + 1 >= rs && rr + 1 < re) P[rr + 1] = y1;
    }
  }
  // column partials: sum over the 8 row-pair lanes, then over the 4 waves
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double x = acc[q].x, yv = acc[q].y;
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
      x += __shfl_xor(x, off, 64);
      yv += __shfl_xor(yv, off, 64);
    }
    if (rp == 0) red[wv][cb + q] = mk(x, yv);
  }
  __syncthreads();
  if (threadIdx.x < TS) {
    z2 t = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
    P2_[(((size_t)k * ntmax + J) * msmax + s) * TS + threadIdx.x] = t;
  }
}

// LAYOUT 0: hemv_strip16; 1: hemv_strip8 (FP32 copy only).  3 waves per
// SIMD (at most 168 VGPRs, no spills): the loads in flight, not the
// arithmetic, bound it.
template <typename TM, int SG, int LAYOUT>
__global__ void __launch_bounds__(256, 3) k_trd_hemv(
    i64 n, i64 j, i64 i0, int jj, const z2* A_, size_t sA, i64 lda, const TM* M_, size_t sM,
    const z2* W_, const double* npart_, z2* P_, z2* P2_, z2* tpart_, z2* scal_, i64 ntmax,
    i64 msmax, i64 nbmax) {
  const int k = blockIdx.z;
  const i64 s = blockIdx.y;
  const z2* A = A_ + k * sA;
  const double* npart = npart_ + (size_t)k * nbmax;
  if (blockIdx.x < TRD_PB) {
    hemv_panel_block<SG>(n, j, i0, jj, A, lda, W_, npart, scal_, tpart_, msmax, k, s);
    return;
  }
  if constexpr (LAYOUT == 1)
    hemv_strip8<SG>(n, j, A, lda, M_ + k * sM, npart, P_, P2_, ntmax, msmax, k, s);
  else
    hemv_strip16<TM, SG>(n, j, A, lda, M_ + k * sM, npart, P_, P2_, ntmax, msmax, k, s);
}

// ---------------------------------------------------------------- K3: finalize w
__global__ void __launch_bounds__(256) k_trd_fin(i64 n, i64 j, i64 i0, int jj, z2* A_, size_t sA,
                                                 i64 lda, z2* W_, double* e_, z2* tau_,
                                                 const z2* scal_, const z2* P_, const z2* P2_,
                                                 const z2* tpart_, z2* dpart_, i64 ntmax, i64 msmax,
                                                 i64 nbmax, int segr) {
  const int k = blockIdx.y;
  const i64 base = j + 1, m = n - base, nt = nt_of(m), ns0 = nseg_of(m, 0, segr);
  (void)nt;
  z2* A = A_ + k * sA;
  z2* W = W_ + (size_t)k * n * NB;
  const z2 tau = scal_[k * 4 + 0], scale = scal_[k * 4 + 1];
  const double beta = scal_[k * 4 + 2].x;
  __shared__ z2 t1[NB], t2[NB];
  __shared__ double sh[256];
  // W^H v and V^H v: the ns0 * PB chunk partials of each of the 2 NB entries,
  // summed by 256 / (2 NB) threads per entry (strided chunks, a short chain
  // each) and then in a fixed order in LDS
  {
    constexpr int NE = 2 * NB, NL = 256 / NE;
    __shared__ z2 tsum[NL][NE];
    const int e = threadIdx.x % NE, c0 = threadIdx.x / NE;
    const z2* tp = tpart_ + (size_t)k * msmax * TRD_PB * 2 * NB;
    z2 a = mk(0.0, 0.0);
    for (i64 s = c0; s < ns0 * TRD_PB; s += NL) a = a + tp[s * 2 * NB + e];
    tsum[c0][e] = a;
    __syncthreads();
    if (threadIdx.x < NE) {
      z2 t = tsum[0][threadIdx.x];
#pragma unroll
      for (int q = 1; q < NL; ++q) t = t + tsum[q][threadIdx.x];
      if (threadIdx.x < NB) t1[threadIdx.x] = t;
      else t2[threadIdx.x - NB] = t;
    }
  }
  __syncthreads();
  // FIN_ROWS rows per block, 4 lanes per row: lane g sums the strip partials
  // J = g, g + 4, ... and panel columns q = g, g + 4, ... (two accumulators
  // each, so several independent loads are in flight), then the 4 lane sums
  // are combined by a fixed butterfly
  const int g = threadIdx.x & 3;
  const i64 r = base + (i64)blockIdx.x * FIN_ROWS + (threadIdx.x >> 2);
  z2 dot = mk(0.0, 0.0);
  const bool live = r < n;
  z2 y = mk(0.0, 0.0);
  i64 rel = 0, R = 0, rr = 0;
  if (live) {
    rel = r - base;
    R = rel / TS;
    rr = rel % TS;
    const z2* P = P_ + (size_t)k * ntmax * n + rel;
    z2 y2 = mk(0.0, 0.0);
    i64 J = g;
    for (; J + 4 <= R; J += 8) {
      y = y + P[(size_t)J * n];
      y2 = y2 + P[(size_t)(J + 4) * n];
    }
    if (J <= R) y = y + P[(size_t)J * n];
    z2 p1 = mk(0.0, 0.0), p2 = mk(0.0, 0.0);
    for (int q = g; q < jj; q += 4) {
      p1 = p1 + A[r + lda * (i0 + q)] * t1[q];
      p2 = p2 + W[r + n * q] * t2[q];
    }
    y = (y + y2) - (p1 + p2);
  }
  y.x += __shfl_xor(y.x, 1, 64);
  y.y += __shfl_xor(y.y, 1, 64);
  y.x += __shfl_xor(y.x, 2, 64);
  y.y += __shfl_xor(y.y, 2, 64);
  if (live && g == 0) {
    const i64 nseg = nseg_of(m, R, segr);
    const z2* P2 = P2_ + ((size_t)k * ntmax + R) * msmax * TS;
    for (i64 s = 0; s < nseg; ++s) y = y + P2[s * TS + rr];
    const z2 v = (r == base) ? mk(1.0, 0.0) : A[r + lda * j] * scale;
    const z2 w = tau * y;
    W[r + n * jj] = w;
    A[r + lda * j] = v;
    dot = cmul(w, v);
  }
  const double dr = block_sum256(dot.x, sh);
  const double di = block_sum256(dot.y, sh);
  if (threadIdx.x == 0) {
    dpart_[(size_t)k * nbmax + blockIdx.x] = mk(dr, di);
    if (blockIdx.x == 0) {
      e_[(size_t)k * n + j] = beta;
      tau_[(size_t)k * n + j] = tau;
    }
  }
}

// ---------------------------------------------------------------- her2k
// A(p:n, p:n) -= V W^H + W V^H on the lower triangle, W = W(p:n, 0:nbp),
// V = V_(p:n, 0:nbp) with leading dimension ldv (the one-stage reduction:
// A(p:n, i0:i0+nbp), whose W's last column still lacks its -tau/2 (w^H v) v
// term, added here while staging when corr is set; the two-stage reduction:
// its panel's reflectors, corr = 0); 64x64 output tiles, lower tiles only.
// On the FP64 matrix cores (v_mfma_f64_16x16x4f64).  With
// X = [Vr Vi Wr Wi] for the tile's rows and, for its columns,
// Yre = [Wr Wi Vr Vi], Yim = [-Wi Wr -Vi Vr]:
//   Re(V W^H + W V^H) = X Yre^T,   Im(V W^H + W V^H) = X Yim^T,
// two real products with k = 4 nbp.  A 64x64 tile per block; each wave owns
// a 32x32 quadrant as 2x2 blocks of 16x16 (real and imaginary accumulators).
// Panel columns are staged 8 at a time as their four real components; the
// k-step of 4 covers panel columns q0 + 4 (lane / 16) ... of one component.
// MFMA 16x16x4 f64 lane layout (checked on the device): A row lane % 16,
// k lane / 16; B k lane / 16, column lane % 16; D row lane / 16 + 4 i,
// column lane % 16 for output i = 0..3.
typedef double f64x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256, 2) k_trd_her2k_mfma(i64 n, i64 p, i64 i0, int nbp, z2* A_,
                                                        size_t sA, i64 lda, float2* S_,
                                                        size_t sS, const z2* W_, const z2* tau_,
                                                        const z2* dpart_, i64 nbmax,
                                                        const z2* V_, size_t sV, i64 ldv, int corr) {
  const int k = blockIdx.y;
  z2* A = A_ + k * sA;
  float2* S = S_ ? S_ + k * sS : nullptr;
  const z2* W = W_ + (size_t)k * n * NB;
  const z2* Vm = V_ + k * sV;
  const i64 jl = i0 + nbp - 1;  // last reflector of the panel
  const z2 alpha =
      corr ? alpha_of(n, jl, tau_[(size_t)k * n + jl], dpart_ + (size_t)k * nbmax) : mk(0.0, 0.0);
  i64 t = blockIdx.x;
  i64 I = (i64)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= t) ++I;
  while (I * (I + 1) / 2 > t) --I;
  const i64 J = t - I * (I + 1) / 2;
  const i64 r0 = p + I * 64, c0 = p + J * 64;
  // [component: Vr Vi Wr Wi][panel column 0..7][tile row / column]
  __shared__ double Xr[4][8][65], Xc[4][8][65];  // padded: the 4 k rows of a step hit different banks
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int qr = wv >> 1, qc = wv & 1;  // the wave's quadrant
  const int l16 = lane & 15, lk = lane >> 4;
  f64x4 re[2][2], im[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) re[u][v] = im[u][v] = f64x4{0.0, 0.0, 0.0, 0.0};
  const z2 zero = mk(0.0, 0.0);
  // the tile elements this lane updates, loaded before the products so their
  // latency is hidden behind them (outside the lower triangle or the matrix:
  // the always-valid element A(p, p), discarded)
  z2 aold[2][2][4];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const i64 r = r0 + qr * 32 + u * 16 + lk + 4 * i, c = c0 + qc * 32 + v * 16 + l16;
        const bool ok = r < n && c < n && r >= c;
        aold[u][v][i] = A[ok ? r + lda * c : p + lda * p];
      }
  // staging of 8 panel columns: this thread's two elements e = tid, tid + 256
  // (panel column e / 64, tile row / column e % 64), loaded into registers
  // one step ahead so the loads overlap the previous step's MFMAs
  const int rr = threadIdx.x & 63, qq0 = threadIdx.x >> 6;  // e = tid: column qq0, e = tid + 256: qq0 + 4
  const bool rin = r0 + rr < n, cin = c0 + rr < n;
  const z2* Ar = Vm + (r0 + rr);
  const z2* Ac = Vm + (c0 + rr);
  const z2* Wr = W + (r0 + rr);
  const z2* Wc = W + (c0 + rr);
  z2 sv[2][4];
#define HER2K_LOAD_STEP(Q0)                                                   \
  _Pragma("unroll") for (int h = 0; h < 2; ++h) {                             \
    const int q = (Q0) + qq0 + 4 * h;                                         \
    const bool qv = q < nbp;                                                  \
    sv[h][0] = (qv && rin) ? Ar[ldv * q] : zero;                              \
    sv[h][1] = (qv && cin) ? Ac[ldv * q] : zero;                              \
    sv[h][2] = (qv && rin) ? Wr[n * q] : zero;                                \
    sv[h][3] = (qv && cin) ? Wc[n * q] : zero;                                \
  }
  HER2K_LOAD_STEP(0)
  for (int q0 = 0; q0 < nbp; q0 += 8) {
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int qq = qq0 + 4 * h, q = q0 + qq;
      const z2 vr = sv[h][0], vc = sv[h][1];
      z2 wr = sv[h][2], wc = sv[h][3];
      if (corr && q == nbp - 1) {
        wr = wr + alpha * vr;
        wc = wc + alpha * vc;
      }
      Xr[0][qq][rr] = vr.x; Xr[1][qq][rr] = vr.y; Xr[2][qq][rr] = wr.x; Xr[3][qq][rr] = wr.y;
      Xc[0][qq][rr] = vc.x; Xc[1][qq][rr] = vc.y; Xc[2][qq][rr] = wc.x; Xc[3][qq][rr] = wc.y;
    }
    __syncthreads();
    if (q0 + 8 < nbp) { HER2K_LOAD_STEP(q0 + 8) }
    // component pairs: X comp -> (Yre comp, sign), (Yim comp, sign)
    //   Vr: (Wr, +) (Wi, -);  Vi: (Wi, +) (Wr, +);  Wr: (Vr, +) (Vi, -);  Wi: (Vi, +) (Vr, +)
#pragma unroll
    for (int comp = 0; comp < 4; ++comp) {
      const int cre = comp ^ 2, cim = (comp ^ 2) ^ 1;
      const double sim = (comp & 1) ? 1.0 : -1.0;
#pragma unroll
      for (int kq = 0; kq < 8; kq += 4) {
        const int q = kq + lk;
        double a[2], bre[2], bim[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) a[u] = Xr[comp][q][qr * 32 + u * 16 + l16];
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          bre[v] = Xc[cre][q][qc * 32 + v * 16 + l16];
          bim[v] = sim * Xc[cim][q][qc * 32 + v * 16 + l16];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int v = 0; v < 2; ++v) {
            re[u][v] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], bre[v], re[u][v], 0, 0, 0);
            im[u][v] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], bim[v], im[u][v], 0, 0, 0);
          }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const i64 r = r0 + qr * 32 + u * 16 + lk + 4 * i, c = c0 + qc * 32 + v * 16 + l16;
        if (r < n && c < n && r >= c) {
          const z2 a = aold[u][v][i] - mk(re[u][v][i], im[u][v][i]);
          A[r + lda * c] = a;
          if (S) S[r + lda * c] = make_float2((float)a.x, (float)a.y);
        }
      }
}
#undef HER2K_LOAD_STEP

// The two-stage reduction's trailing update: A(p:n, p:n) -= V W^H + W V^H with
// V its panel's reflectors (w.ts->V, leading dimension n) and W in w.W.
int her2k_twostage(TridiagWork& w, i64 p, i64 i0, int nbp, hipStream_t s) {
  const i64 mt = w.n - p, ntt = (mt + 63) / 64;
  if (mt <= 0 || nbp <= 0) return 0;
  hipLaunchKernelGGL(k_trd_her2k_mfma, dim3((unsigned)(ntt * (ntt + 1) / 2), w.K), dim3(256), 0, s,
                     w.n, p, i0, nbp, w.A, w.sA, w.lda, (float2*)nullptr, (size_t)0, w.W, w.tau,
                     w.dpart, npart_stride(w.n), w.ts->V, (size_t)w.n * TS_KD, w.n, 0);
  FORGE_LAUNCH();
  return 0;
}

// S = fp32(A) on the lower triangle (column c, rows c..n-1); grid
// (cdiv(n, 256), ny, K), columns blockIdx.y, blockIdx.y + ny, ...
__global__ void __launch_bounds__(256) k_trd_to_fp32(i64 n, const z2* A_, size_t sA, i64 lda,
                                                     float2* S_, size_t sS) {
  for (i64 c = blockIdx.y; c < n; c += gridDim.y) {
    const i64 r = c + (i64)blockIdx.x * 256 + threadIdx.x;
    if (r < n) {
      const z2 a = A_[blockIdx.z * sA + r + lda * c];
      S_[blockIdx.z * sS + r + lda * c] = make_float2((float)a.x, (float)a.y);
    }
  }
}

__global__ void k_trd_clear(size_t count, z2* __restrict__ p) {
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < count; q += (size_t)gridDim.x * 256)
    p[q] = mk(0.0, 0.0);
}

__global__ void k_trd_lastdiag(i64 n, const z2* A_, size_t sA, i64 lda, double* d_) {
  int k = blockIdx.x;
  if (threadIdx.x == 0) d_[(size_t)k * n + n - 1] = A_[k * sA + (n - 1) + lda * (n - 1)].x;
}

// ---------------------------------------------------------------- driver
// Timing (hemv_ms != nullptr): the hemv launches of every TIME_STRIDE-th
// column carry a pair of HIP events (hipExtLaunchKernel), so the events take
// the dispatch's own begin and end timestamps; their summed time, count and
// algorithmic bytes (the lower triangle of the m x m trailing matrix read
// once: 8 m^2 per matrix in FP64, 4 m^2 from the FP32 copy) are accumulated
// into the outputs.  The pairs form a ring
// of TIME_RING entries; a pair is read back before it is reused, which waits
// only for a launch TIME_RING * TIME_STRIDE columns back.  The stride is
// coprime with the panel width, so the sample visits every panel position
// (the hemv's panel term grows with the position in the panel).
constexpr int TIME_STRIDE = 17;
constexpr int TIME_RING = 8;

static int read_pair(TridiagWork& w, i64 q, double* ms) {
  float ms_ = 0.f;
  FORGE_HIP(hipEventSynchronize(w.evs[2 * q + 1]));
  FORGE_HIP(hipEventElapsedTime(&ms_, w.evs[2 * q], w.evs[2 * q + 1]));
  *ms += ms_;
  return 0;
}

// hemv launch; with events e0/e1 they take the dispatch's begin and end
template <typename TM, int SG, int LAYOUT>
static void launch_hemv_v(dim3 g, hipStream_t s, hipEvent_t e0, hipEvent_t e1, i64 n, i64 j,
                          i64 i0, int jj, const TridiagWork& w, const TM* M, size_t sM, i64 ntmax,
                          i64 msmax, i64 nbmax) {
  if (e0)
    hipExtLaunchKernelGGL(k_trd_hemv<TM, SG, LAYOUT>, g, dim3(256), 0, s, e0, e1, 0, n, j, i0, jj,
                          (const z2*)w.A, w.sA, w.lda, M, sM, (const z2*)w.W,
                          (const double*)w.npart, w.P, w.P2, w.tpart, w.scal, ntmax, msmax, nbmax);
  else
    hipLaunchKernelGGL((k_trd_hemv<TM, SG, LAYOUT>), g, dim3(256), 0, s, n, j, i0, jj,
                       (const z2*)w.A, w.sA, w.lda, M, sM, (const z2*)w.W, (const double*)w.npart,
                       w.P, w.P2, w.tpart, w.scal, ntmax, msmax, nbmax);
}

// a hemv variant: segment height and lane layout
struct HemvVariant {
  int sg, layout;
};

// the hemv of column j over the first Kj matrices with variant v
static void launch_hemv(HemvVariant v, hipStream_t s, hipEvent_t e0, hipEvent_t e1, i64 j, i64 i0,
                        int jj, int Kj, const TridiagWork& w) {
  const i64 n = w.n, m = n - j - 1, nt = nt_of(m);
  const i64 ntmax = nt_of(n), msmax = nseg_of(n, 0), nbmax = npart_stride(n);
  const dim3 g((unsigned)(nt + TRD_PB), (unsigned)nseg_of(m, 0, v.sg), Kj);
  if (w.S) {
    const float2* S = (const float2*)w.S;
    if (v.layout == 1) {
      if (v.sg == 1024) launch_hemv_v<float2, 1024, 1>(g, s, e0, e1, n, j, i0, jj, w, S, w.sS, ntmax, msmax, nbmax);
      else launch_hemv_v<float2, 2048, 1>(g, s, e0, e1, n, j, i0, jj, w, S, w.sS, ntmax, msmax, nbmax);
    } else {
      if (v.sg == 1024) launch_hemv_v<float2, 1024, 0>(g, s, e0, e1, n, j, i0, jj, w, S, w.sS, ntmax, msmax, nbmax);
      else launch_hemv_v<float2, 2048, 0>(g, s, e0, e1, n, j, i0, jj, w, S, w.sS, ntmax, msmax, nbmax);
    }
  } else {
    const z2* A = (const z2*)w.A;
    if (v.sg == 1024) launch_hemv_v<z2, 1024, 0>(g, s, e0, e1, n, j, i0, jj, w, A, w.sA, ntmax, msmax, nbmax);
    else launch_hemv_v<z2, 2048, 0>(g, s, e0, e1, n, j, i0, jj, w, A, w.sA, ntmax, msmax, nbmax);
  }
}

// ---------------------------------------------------------------- hemv tuner
// The hemv variant per band of trailing orders, measured on first use for a
// shape (device, n, K, FP64 or FP32 copy) and kept for the process:
//   * the segment height (rows per strip block, 1024 or 2048): tall segments
//     give each block more bytes to stream and fewer partials, short ones
//     give more blocks, which fills the chip when the trailing matrix is
//     small;
//   * for the FP32 copy, the lane layout: 16 lanes per 16-row column chunk
//     with 8-byte loads, or 8 lanes with 16-byte row-pair loads.
// In each band the order reaches, the hemv of one column at a representative
// trailing order (the band's middle, at most n - 1) is launched three times
// with each candidate and the last two are timed with HIP events.  The hemv
// writes only its own outputs (partials, reflector scalars), which the
// reduction rewrites before it reads them.  FORGE_TUNE_LOG=1 prints the
// timings.
constexpr int HEMV_BANDS = 4;
constexpr i64 HEMV_BAND_LO[HEMV_BANDS] = {0, 1536, 3072, 6144};
constexpr i64 HEMV_BAND_HI[HEMV_BANDS] = {1536, 3072, 6144, (i64)1 << 40};
constexpr HemvVariant HEMV_CAND[4] = {{1024, 0}, {2048, 0}, {1024, 1}, {2048, 1}};

struct HemvPlan {
  HemvVariant v[HEMV_BANDS];
};

static std::mutex g_hemv_mu;
static std::map<std::tuple<int, i64, int, int>, HemvPlan> g_hemv_plan;

static int hemv_band(i64 m) {
  int b = 0;
  while (b + 1 < HEMV_BANDS && m >= HEMV_BAND_LO[b + 1]) ++b;
  return b;
}

static HemvPlan default_plan(bool fp32) {
  HemvPlan p;
  for (int b = 0; b < HEMV_BANDS; ++b) p.v[b] = fp32 ? HemvVariant{2048, 0} : HemvVariant{1024, 0};
  return p;
}

static std::tuple<int, i64, int, int> plan_key(const TridiagWork& w) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  return std::make_tuple(dev, w.n, w.K, w.S ? 1 : 0);
}

// FORGE_HEMV_FORCE="<segment>,<layout>" (1024 or 2048; 0 or 1, the latter for
// the FP32 copy only) replaces the tuned plan by that one variant in every
// band, for tests of each variant; 0 when unset or not applicable.
static int forced_plan(bool fp32, HemvPlan* out) {
  const char* f = getenv("FORGE_HEMV_FORCE");
  if (!f) return 0;
  int sg = 0, layout = -1;
  if (sscanf(f, "%d,%d", &sg, &layout) != 2 || (sg != 1024 && sg != 2048) || layout < 0 ||
      layout > 1 || (!fp32 && layout != 0))
    return 0;
  for (int b = 0; b < HEMV_BANDS; ++b) out->v[b] = HemvVariant{sg, layout};
  return 1;
}

// the plan of w's shape; tuned here when absent and may_tune (never inside a
// stream capture), else the default variants
static int hemv_plan(TridiagWork& w, hipStream_t s, bool may_tune, HemvPlan* out) {
  if (forced_plan(w.S != nullptr, out)) return 0;
  const auto key = plan_key(w);
  {
    std::lock_guard<std::mutex> g(g_hemv_mu);
    auto it = g_hemv_plan.find(key);
    if (it != g_hemv_plan.end()) { *out = it->second; return 0; }
  }
  const bool fp32 = w.S != nullptr;
  HemvPlan p = default_plan(fp32);
  if (!may_tune) { *out = p; return 0; }
  const i64 n = w.n, nbmax = npart_stride(n);
  const bool log = getenv("FORGE_TUNE_LOG") != nullptr;
  // the norm partials the reflector reads: cleared, so the timed launches
  // compute with finite numbers
  hipLaunchKernelGGL(k_trd_clear, dim3(64), dim3(256), 0, s, (size_t)w.K * nbmax / 2, (z2*)w.npart);
  FORGE_LAUNCH();
  hipEvent_t e0, e1;
  FORGE_HIP(hipEventCreate(&e0));
  FORGE_HIP(hipEventCreate(&e1));
  const int ncand = fp32 ? 4 : 2;
  for (int b = 0; b < HEMV_BANDS; ++b) {
    if (n - 1 <= HEMV_BAND_LO[b] || n - 1 < 64) continue;
    const i64 hi = std::min<i64>(HEMV_BAND_HI[b], n - 1);
    const i64 m = (HEMV_BAND_LO[b] + hi) / 2;
    if (m < 1) continue;
    const i64 j = n - 1 - m;
    float best = 0.f;
    for (int c = 0; c < ncand; ++c) {
      const HemvVariant v = HEMV_CAND[c];
      launch_hemv(v, s, nullptr, nullptr, j, j, 0, w.K, w);
      FORGE_LAUNCH();
      FORGE_HIP(hipEventRecord(e0, s));
      launch_hemv(v, s, nullptr, nullptr, j, j, 0, w.K, w);
      launch_hemv(v, s, nullptr, nullptr, j, j, 0, w.K, w);
      FORGE_LAUNCH();
      FORGE_HIP(hipEventRecord(e1, s));
      FORGE_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      FORGE_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (log) {
        const double gbs = (fp32 ? 4.0 : 8.0) * (double)m * (double)m * w.K * 2.0 / (ms * 1e-3) / 1e9;
        fprintf(stderr, "[hemv tuner] n=%lld K=%d %s m=%lld segment %d layout %d: %.1f us, %.0f GB/s\n",
                (long long)n, w.K, fp32 ? "fp32" : "fp64", (long long)m, v.sg, v.layout,
                ms * 500.f, gbs);
      }
      if (c == 0 || ms < best) {
        best = ms;
        p.v[b] = v;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::lock_guard<std::mutex> g(g_hemv_mu);
  g_hemv_plan[key] = p;
  *out = p;
  return 0;
}

int tridiag_hemv_plan(const TridiagWork& w, int* code_by_band, i64* band_lo) {
  HemvPlan forced;
  const bool fp32 = w.S != nullptr;
  if (forced_plan(fp32, &forced)) {
    for (int b = 0; b < HEMV_BANDS; ++b) {
      code_by_band[b] = forced.v[b].sg / 1024 + 2 * forced.v[b].layout;
      band_lo[b] = HEMV_BAND_LO[b];
    }
    return 1;
  }
  const auto key = plan_key(w);
  std::lock_guard<std::mutex> g(g_hemv_mu);
  auto it = g_hemv_plan.find(key);
  const HemvPlan p = it != g_hemv_plan.end() ? it->second : default_plan(fp32);
  for (int b = 0; b < HEMV_BANDS; ++b) {
    code_by_band[b] = p.v[b].sg / 1024 + 2 * p.v[b].layout;
    band_lo[b] = HEMV_BAND_LO[b];
  }
  return it != g_hemv_plan.end() ? 1 : 0;
}

// The launch sequence of one reduction.  capturing: inside a stream capture
// (no host synchronisation).
static int enqueue_reduction(TridiagWork& w, hipStream_t s, double* hemv_ms, long* hemv_launches,
                             double* hemv_bytes, bool capturing) {
  const i64 n = w.n;
  const int K = w.K;
  w.two_stage = 0;
  w.voff = 1;
  if (w.ts && !w.first_col) return twostage_reduce(w, s, capturing);
  const i64 ntmax = nt_of(n), msmax = nseg_of(n, 0), nbmax = npart_stride(n);
  const bool timed = hemv_ms != nullptr;
  if (timed && w.nev == 0) {
    w.evs = new hipEvent_t[2 * TIME_RING];
    for (int q = 0; q < 2 * TIME_RING; ++q) FORGE_HIP(hipEventCreate(&w.evs[q]));
    w.nev = 2 * TIME_RING;
  }
  i64 nrec = 0;
  double bytes = 0.0, tms = 0.0;
  // uneven sizes: a matrix joins at its first column with W, tau and the w^H v
  // partials of the columns before it zero, so those columns act as identity
  // reflectors in the panel's column updates, panel products and her2k (their
  // Householder vectors lie in the slot's zero columns)
  int Kj = K;
  if (w.first_col) {
    for (int k = 1; k < K; ++k)
      if (w.first_col[k] < w.first_col[k - 1] || w.first_col[k] < 0 || w.first_col[k] >= n) {
        set_error("tridiag: first_col must be non-decreasing and below n");
        return -2;
      }
    const size_t cW = (size_t)K * n * NB, cT = (size_t)K * n, cD = (size_t)K * nbmax;
    hipLaunchKernelGGL(k_trd_clear, dim3(1024), dim3(256), 0, s, cW, w.W);
    hipLaunchKernelGGL(k_trd_clear, dim3(256), dim3(256), 0, s, cT, w.tau);
    hipLaunchKernelGGL(k_trd_clear, dim3(64), dim3(256), 0, s, cD, w.dpart);
    FORGE_LAUNCH();
    Kj = 0;
  }
  HemvPlan plan;
  {
    const int r = hemv_plan(w, s, !capturing, &plan);
    if (r) return r;
  }
  if (w.S) {
    hipLaunchKernelGGL(k_trd_to_fp32, dim3((unsigned)cdiv(n, 256), (unsigned)std::min<i64>(n, 512), K),
                       dim3(256), 0, s, n, w.A, w.sA, w.lda, w.S, w.sS);
    FORGE_LAUNCH();
  }
  for (i64 i0 = 0; i0 < n - 1; i0 += NB) {
    const int nbp = (int)((n - 1 - i0) < NB ? (n - 1 - i0) : NB);
    for (int jj = 0; jj < nbp; ++jj) {
      const i64 j = i0 + jj;
      if (w.first_col)
        while (Kj < K && w.first_col[Kj] <= j) ++Kj;
      if (Kj == 0) continue;
      hipLaunchKernelGGL(k_trd_col, dim3(cdiv(n - j, 256), Kj), dim3(256), 0, s, n, j, i0, jj, w.A,
                         w.sA, w.lda, w.W, w.d, w.tau, w.npart, w.dpart, nbmax);
      FORGE_LAUNCH();
      const i64 m = n - j - 1;
      const HemvVariant hv = plan.v[hemv_band(m)];
      const int segr = hv.sg;
      const bool tj = timed && (j % TIME_STRIDE == 0);
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (tj) {
        const i64 q = nrec % TIME_RING;
        if (nrec >= TIME_RING && read_pair(w, q, &tms)) return -1;
        e0 = w.evs[2 * q];
        e1 = w.evs[2 * q + 1];
        ++nrec;
        // algorithmic bytes: the lower triangle of the m x m trailing matrix
        bytes += (w.S ? 4.0 : 8.0) * (double)m * (double)m * Kj;
      }
      launch_hemv(hv, s, e0, e1, j, i0, jj, Kj, w);
      FORGE_LAUNCH();
      hipLaunchKernelGGL(k_trd_fin, dim3((unsigned)nfin_of(m), Kj), dim3(256), 0, s, n, j, i0, jj, w.A,
                         w.sA, w.lda, w.W, w.e, w.tau, w.scal, w.P, w.P2, w.tpart, w.dpart, ntmax,
                         msmax, nbmax, segr);
      FORGE_LAUNCH();
    }
    const i64 p = i0 + nbp, mt = n - p, ntt = (mt + 63) / 64;
    if (Kj == 0) continue;
    hipLaunchKernelGGL(k_trd_her2k_mfma, dim3((unsigned)(ntt * (ntt + 1) / 2), Kj), dim3(256), 0, s,
                       n, p, i0, nbp, w.A, w.sA, w.lda, w.S, w.sS, w.W, w.tau, w.dpart, nbmax,
                       w.A + w.lda * i0, w.sA, w.lda, 1);
    FORGE_LAUNCH();
    // bound the launches in flight (about 400): the host otherwise runs up to
    // 3 n launches ahead, which rocprofv3's dispatch tracing does not survive
    // (host fault in its record buffers beyond ~15000 pending dispatches);
    // the GPU idles only for the next launch's latency, ~100 times per call
    if (!capturing && (i0 / NB) % 4 == 3) FORGE_HIP(hipStreamSynchronize(s));
  }
  hipLaunchKernelGGL(k_trd_lastdiag, dim3(K), dim3(64), 0, s, n, w.A, w.sA, w.lda, w.d);
  FORGE_LAUNCH();
  if (timed) {
    for (i64 r = (nrec > TIME_RING ? nrec - TIME_RING : 0); r < nrec; ++r)
      if (read_pair(w, r % TIME_RING, &tms)) return -1;
    *hemv_ms += tms;
    *hemv_launches += nrec;
    *hemv_bytes += bytes;
  }
  return 0;
}

// A captured reduction and the shape it was captured for: the kernel
// arguments are fixed by n, K, the active prefix of every column
// (first_col), the FP32 copy and the buffer addresses.
struct TrdGraph {
  hipGraphExec_t exec = nullptr;
  i64 n = 0, lda = 0;
  int K = 0;
  const void *A = nullptr, *S = nullptr, *W = nullptr, *ts = nullptr;
  size_t sA = 0, sS = 0;
  std::vector<i64> first;
  bool same(const TridiagWork& w) const {
    if (n != w.n || K != w.K || lda != w.lda || A != w.A || S != w.S || W != w.W || ts != w.ts || sA != w.sA ||
        sS != w.sS || first.size() != (w.first_col ? (size_t)w.K : 0))
      return false;
    for (size_t q = 0; q < first.size(); ++q)
      if (first[q] != w.first_col[q]) return false;
    return true;
  }
};

void tridiag_release(TridiagWork& w) {
  auto* g = (TrdGraph*)w.graph;
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  delete g;
  w.graph = nullptr;
}

int tridiag_enqueue(TridiagWork& w, hipStream_t s) {
  return enqueue_reduction(w, s, nullptr, nullptr, nullptr, true);
}

int tridiag(TridiagWork& w, hipStream_t s, double* hemv_ms, long* hemv_launches,
            double* hemv_bytes) {
  if (!w.use_graph || hemv_ms) return enqueue_reduction(w, s, hemv_ms, hemv_launches, hemv_bytes, false);
  auto* g = (TrdGraph*)w.graph;
  if (!g || !g->same(w)) {
    tridiag_release(w);
    if (!w.ts || w.first_col) {
      // the hemv plan is measured before the capture (it times with events)
      HemvPlan plan;
      const int r = hemv_plan(w, s, true, &plan);
      if (r) return r;
    }
    hipGraph_t graph = nullptr;
    FORGE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    const int rc = enqueue_reduction(w, s, nullptr, nullptr, nullptr, true);
    const hipError_t ce = hipStreamEndCapture(s, &graph);
    if (rc != 0) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (ce != hipSuccess) {
      set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
      return -1;
    }
    g = new TrdGraph();
    const hipError_t ie = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      delete g;
      set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
      return -1;
    }
    g->n = w.n; g->K = w.K; g->lda = w.lda; g->A = w.A; g->S = w.S; g->W = w.W; g->ts = w.ts;
    g->sA = w.sA; g->sS = w.sS;
    if (w.first_col) g->first.assign(w.first_col, w.first_col + w.K);
    w.graph = g;
  }
  FORGE_HIP(hipGraphLaunch(g->exec, s));
  return 0;
}

}  // namespace forge
