// Batched reduction of K Hermitian matrices (lower triangle, column-major) to
// real symmetric tridiagonal form A = Q T Q^H, Q = H_0 H_1 ... H_{n-2},
// H_j = I - tau_j v_j v_j^H.  This is the blocked lower Householder reduction
// that LAPACK zhetrd/zlatrd define (the reference reaches it through zheevr,
// LapackRoutines.f90:414-426); the arithmetic per column is the same:
//
//   column update   a(j:n) -= A(j:n, i0:j) conj(W(j, :)) + W(j:n, :) conj(A(j, i0:j))
//   reflector       beta = -sign(Re alpha) |(alpha, x)|, tau = (beta - alpha)/beta,
//                   v = (1, x / (alpha - beta)), e_j = beta
//   w               = tau (A22 v - V (W^H v) - W (V^H v)),   w += -tau/2 (w^H v) v
//   panel end       A22 -= V W^H + W V^H   (lower triangle)
//
// Each column is three launches over all K matrices:
//   k_trd_col    column update; the last block to finish forms the reflector.
//   k_trd_hemv   A22 v over lower-triangle tiles (A22 read once, LDS-staged 32x32
//                tiles, each tile feeding both A v and A^H v), plus W^H v, V^H v.
//   k_trd_fin    reduces the hemv partials into w; the last block applies the
//                -tau/2 (w^H v) v correction.
// Completion counters implement the last-block reductions; each last block
// resets its counter, so the counters stay zero between launches.
#include "kernels.hpp"

namespace forge {

constexpr int NB = TRD_NB;
constexpr int TS = TRD_TS;
constexpr int SEG = TRD_SEG;

__host__ __device__ static inline i64 nt_of(i64 m) { return (m + TS - 1) / TS; }
__host__ __device__ static inline i64 seg_of(i64 nt) { return (nt + SEG - 1) / SEG; }
__host__ __device__ static inline i64 nblk_of(i64 m) { return (m + 255) / 256; }

struct TrdLayout {
  size_t A_d, A_e, A_tau, A_W, A_P, A_P2, A_tp, A_t12, A_np, A_scal, A_cnt, total;
};

static TrdLayout layout(i64 n, int K) {
  i64 nt = nt_of(n), ms = seg_of(nt), nb = nblk_of(n) + 1;
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
  L.A_P = take(sizeof(z2) * (size_t)K * nt * nt * TS);
  L.A_P2 = take(sizeof(z2) * (size_t)K * nt * ms * TS);
  L.A_tp = take(sizeof(z2) * (size_t)K * nt * 2 * NB);
  L.A_t12 = take(sizeof(z2) * (size_t)K * 2 * NB);
  L.A_np = take(sizeof(double) * 2 * (size_t)K * nb);
  L.A_scal = take(sizeof(z2) * K * 4);
  L.A_cnt = take(sizeof(unsigned) * K * 4);
  L.total = off;
  return L;
}

size_t tridiag_workspace_bytes(i64 n, int K) { return layout(n, K).total; }

int tridiag_bind(TridiagWork& w, void* base, i64 n, int K) {
  TrdLayout L = layout(n, K);
  char* b = (char*)base;
  w.n = n;
  w.K = K;
  w.d = (double*)(b + L.A_d);
  w.e = (double*)(b + L.A_e);
  w.tau = (z2*)(b + L.A_tau);
  w.W = (z2*)(b + L.A_W);
  w.P = (z2*)(b + L.A_P);
  w.P2 = (z2*)(b + L.A_P2);
  w.tpart = (z2*)(b + L.A_tp);
  w.t12 = (z2*)(b + L.A_t12);
  w.npart = (double*)(b + L.A_np);
  w.scal = (z2*)(b + L.A_scal);
  w.cnt = (unsigned*)(b + L.A_cnt);
  w.evs = nullptr;
  w.nev = 0;
  FORGE_HIP(hipMemset(w.cnt, 0, sizeof(unsigned) * K * 4));
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

__device__ inline z2 vload(const z2* p) {
  const volatile double* q = (const volatile double*)p;
  return z2{q[0], q[1]};
}
__device__ inline double vloadd(const double* p) { return *(const volatile double*)p; }

// Returns true in every thread of the last block (of gridDim.x blocks on this
// y/z slice) to arrive; that block then sees all other blocks' writes.
__device__ inline bool last_block(unsigned* counter, unsigned total) {
  __shared__ int is_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned prev = atomicAdd(counter, 1u);
    is_last = (prev == total - 1);
  }
  __syncthreads();
  if (is_last) __threadfence();
  return is_last != 0;
}

// ---------------------------------------------------------------- K1: column update
__global__ void __launch_bounds__(256) k_trd_col(i64 n, i64 j, i64 i0, int jj, z2* A_,
                                                 size_t sA, z2* W_, double* d_, double* e_,
                                                 z2* tau_, double* npart_, z2* scal_,
                                                 unsigned* cnt_, i64 nbmax) {
  const int k = blockIdx.y;
  z2* A = A_ + k * sA;
  const z2* W = W_ + (size_t)k * n * NB;
  double* npart = npart_ + (size_t)k * 2 * nbmax;
  __shared__ z2 arow[NB], wrow[NB];
  __shared__ double sh[256];
  if (threadIdx.x < jj) {
    arow[threadIdx.x] = A[j + n * (i0 + threadIdx.x)];
    wrow[threadIdx.x] = W[j + n * threadIdx.x];
  }
  __syncthreads();
  i64 r = j + (i64)blockIdx.x * 256 + threadIdx.x;
  double nrm = 0.0;
  if (r < n) {
    z2 a = A[r + n * j];
    if (r == j) a.y = 0.0;
    for (int q = 0; q < jj; ++q) {
      a = a - mulc(A[r + n * (i0 + q)], wrow[q]);
      a = a - mulc(W[r + n * q], arow[q]);
    }
    if (r == j) {
      a.y = 0.0;
      d_[(size_t)k * n + j] = a.x;
    }
    A[r + n * j] = a;
    if (r >= j + 2) nrm = abs2(a);
  }
  nrm = block_sum256(nrm, sh);
  if (threadIdx.x == 0) npart[blockIdx.x] = nrm;
  if (!last_block(&cnt_[k * 4 + 0], gridDim.x)) return;
  if (threadIdx.x == 0) {
    double xn2 = 0.0;
    for (unsigned b = 0; b < gridDim.x; ++b) xn2 += vloadd(&npart[b]);
    z2 alpha = vload(&A[(j + 1) + n * j]);
    z2 tau, scale;
    double beta;
    if (xn2 == 0.0 && alpha.y == 0.0) {
      tau = mk(0.0, 0.0);
      beta = alpha.x;
      scale = mk(0.0, 0.0);
    } else {
      double nr = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
      beta = alpha.x >= 0.0 ? -nr : nr;
      tau = mk((beta - alpha.x) / beta, -alpha.y / beta);
      z2 den = mk(alpha.x - beta, alpha.y);  // 1 / (alpha - beta)
      double dd = abs2(den);
      scale = mk(den.x / dd, -den.y / dd);
    }
    z2* sc = scal_ + k * 4;
    sc[0] = tau;
    sc[1] = mk(beta, 0.0);
    sc[2] = scale;
    e_[(size_t)k * n + j] = beta;
    tau_[(size_t)k * n + j] = tau;
    cnt_[k * 4 + 0] = 0;
  }
}

__device__ inline z2 vcomp(const z2* A, i64 n, i64 j, i64 base, i64 r, z2 scale) {
  if (r >= n) return mk(0.0, 0.0);
  if (r == base) return mk(1.0, 0.0);
  return A[r + n * j] * scale;
}

// ---------------------------------------------------------------- K2: hemv
// grid (nt, maxseg, K); block (J, s) covers tiles (I, J), I in [J + s*SEG, ...).
__global__ void __launch_bounds__(256) k_trd_hemv(i64 n, i64 j, i64 i0, int jj, const z2* A_,
                                                  size_t sA, const z2* W_, const z2* scal_,
                                                  z2* P_, z2* P2_, z2* tpart_, z2* t12_,
                                                  unsigned* cnt_, i64 ntmax, i64 msmax) {
  const int k = blockIdx.z;
  const i64 J = blockIdx.x, s = blockIdx.y;
  const i64 base = j + 1, m = n - base, nt = nt_of(m), ms = seg_of(nt);
  const i64 Ib = J + s * SEG;
  if (J >= nt || Ib >= nt) return;
  const i64 Ie = Ib + SEG < nt ? Ib + SEG : nt;
  const z2* A = A_ + k * sA;
  const z2 scale = scal_[k * 4 + 2];
  __shared__ z2 tile[TS][TS + 1];
  __shared__ z2 vJ[TS], vI[TS];
  __shared__ z2 red[8][TS];
  const int tx = threadIdx.x & 31, g = threadIdx.x >> 5;
  const i64 c0 = base + J * TS;
  if (threadIdx.x < TS) vJ[threadIdx.x] = vcomp(A, n, j, base, c0 + threadIdx.x, scale);
  z2 yJ = mk(0.0, 0.0);
  z2* P = P_ + (size_t)k * ntmax * ntmax * TS;
  for (i64 I = Ib; I < Ie; ++I) {
    const i64 r0 = base + I * TS;
    __syncthreads();
    if (threadIdx.x < TS) vI[threadIdx.x] = vcomp(A, n, j, base, r0 + threadIdx.x, scale);
    // load tile rows r0.., cols c0.. (lower part only on the diagonal tile)
    for (int q = 0; q < 4; ++q) {
      int cc = g * 4 + q;
      i64 r = r0 + tx, c = c0 + cc;
      z2 a = mk(0.0, 0.0);
      if (r < n && c < n && (I != J || tx >= cc)) a = A[r + n * c];
      tile[tx][cc] = a;
    }
    __syncthreads();
    if (I == J) {  // Hermitian completion of the diagonal tile
      for (int q = 0; q < 4; ++q) {
        int cc = g * 4 + q;
        if (tx < cc) tile[tx][cc] = conjz(tile[cc][tx]);
      }
      __syncthreads();
    }
    // y_I(r) partial over this thread's 4 columns
    z2 acc = mk(0.0, 0.0);
    for (int q = 0; q < 4; ++q) acc = acc + tile[tx][g * 4 + q] * vJ[g * 4 + q];
    red[g][tx] = acc;
    // y_J(c) partial: A^H v_I over this thread's 4 rows (off-diagonal tiles)
    if (I != J) {
      for (int q = 0; q < 4; ++q) yJ = yJ + cmul(tile[g * 4 + q][tx], vI[g * 4 + q]);
    }
    __syncthreads();
    if (threadIdx.x < TS) {
      z2 sacc = red[0][threadIdx.x];
      for (int q = 1; q < 8; ++q) sacc = sacc + red[q][threadIdx.x];
      P[((size_t)I * ntmax + J) * TS + threadIdx.x] = sacc;
    }
  }
  __syncthreads();
  red[g][tx] = yJ;
  __syncthreads();
  if (threadIdx.x < TS) {
    z2 sacc = red[0][threadIdx.x];
    for (int q = 1; q < 8; ++q) sacc = sacc + red[q][threadIdx.x];
    P2_[(((size_t)k * ntmax + J) * msmax + s) * TS + threadIdx.x] = sacc;
  }
  if (s != 0) return;
  // W^H v and V^H v over the rows of tile J (q < jj)
  __syncthreads();
  const z2* W = W_ + (size_t)k * n * NB;
  z2* tpart = tpart_ + (size_t)k * ntmax * 2 * NB;
  {
    int q = threadIdx.x >> 3, part = threadIdx.x & 7;  // 32 columns x 8 row groups
    z2 t1 = mk(0.0, 0.0), t2 = mk(0.0, 0.0);
    if (q < jj) {
      for (int rr = part * 4; rr < part * 4 + 4; ++rr) {
        i64 r = c0 + rr;
        if (r < n) {
          t1 = t1 + cmul(W[r + n * q], vJ[rr]);
          t2 = t2 + cmul(A[r + n * (i0 + q)], vJ[rr]);
        }
      }
    }
    __shared__ z2 tr1[32][8], tr2[32][8];
    tr1[q][part] = t1;
    tr2[q][part] = t2;
    __syncthreads();
    if (threadIdx.x < NB) {
      z2 a = mk(0.0, 0.0), b = mk(0.0, 0.0);
      for (int p = 0; p < 8; ++p) {
        a = a + tr1[threadIdx.x][p];
        b = b + tr2[threadIdx.x][p];
      }
      tpart[J * 2 * NB + threadIdx.x] = a;
      tpart[J * 2 * NB + NB + threadIdx.x] = b;
    }
  }
  // last of the nt segment-0 blocks reduces tpart -> t12
  __shared__ int is_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) is_last = (atomicAdd(&cnt_[k * 4 + 1], 1u) == (unsigned)nt - 1);
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  if (threadIdx.x < 2 * NB) {
    z2 a = mk(0.0, 0.0);
    for (i64 JJ = 0; JJ < nt; ++JJ) a = a + vload(&tpart[JJ * 2 * NB + threadIdx.x]);
    t12_[k * 2 * NB + threadIdx.x] = a;
  }
  if (threadIdx.x == 0) cnt_[k * 4 + 1] = 0;
}

// ---------------------------------------------------------------- K3: finalize w
__global__ void __launch_bounds__(256) k_trd_fin(i64 n, i64 j, i64 i0, int jj, z2* A_, size_t sA,
                                                 z2* W_, const z2* scal_, const z2* P_,
                                                 const z2* P2_, const z2* t12_, double* npart_,
                                                 unsigned* cnt_, i64 ntmax, i64 msmax,
                                                 i64 nbmax) {
  const int k = blockIdx.y;
  const i64 base = j + 1, m = n - base, nt = nt_of(m);
  z2* A = A_ + k * sA;
  z2* W = W_ + (size_t)k * n * NB;
  const z2 tau = scal_[k * 4 + 0], scale = scal_[k * 4 + 2];
  __shared__ z2 t1[NB], t2[NB];
  __shared__ double sh[256];
  if (threadIdx.x < jj) {
    t1[threadIdx.x] = t12_[k * 2 * NB + threadIdx.x];
    t2[threadIdx.x] = t12_[k * 2 * NB + NB + threadIdx.x];
  }
  __syncthreads();
  const i64 r = base + (i64)blockIdx.x * 256 + threadIdx.x;
  z2 dot = mk(0.0, 0.0);
  if (r < n) {
    const i64 R = (r - base) / TS, rr = (r - base) % TS;
    const z2* P = P_ + (size_t)k * ntmax * ntmax * TS;
    z2 y = mk(0.0, 0.0);
    for (i64 J = 0; J <= R; ++J) y = y + P[((size_t)R * ntmax + J) * TS + rr];
    i64 nseg = seg_of(nt - R);
    const z2* P2 = P2_ + ((size_t)k * ntmax + R) * msmax * TS;
    for (i64 s = 0; s < nseg; ++s) y = y + P2[s * TS + rr];
    z2 v = (r == base) ? mk(1.0, 0.0) : A[r + n * j] * scale;
    for (int q = 0; q < jj; ++q) {
      y = y - A[r + n * (i0 + q)] * t1[q];
      y = y - W[r + n * q] * t2[q];
    }
    z2 w = tau * y;
    W[r + n * jj] = w;
    A[r + n * j] = v;
    dot = cmul(w, v);
  }
  double* np = npart_ + (size_t)k * 2 * nbmax;
  double dr = block_sum256(dot.x, sh);
  double di = block_sum256(dot.y, sh);
  if (threadIdx.x == 0) {
    np[2 * blockIdx.x] = dr;
    np[2 * blockIdx.x + 1] = di;
  }
  if (!last_block(&cnt_[k * 4 + 2], gridDim.x)) return;
  __shared__ z2 alpha_s;
  if (threadIdx.x == 0) {
    z2 s = mk(0.0, 0.0);
    for (unsigned b = 0; b < gridDim.x; ++b) s = s + mk(vloadd(&np[2 * b]), vloadd(&np[2 * b + 1]));
    alpha_s = (-0.5) * (tau * s);
    cnt_[k * 4 + 2] = 0;
  }
  __syncthreads();
  const z2 alpha = alpha_s;
  for (i64 rr = base + threadIdx.x; rr < n; rr += 256) {
    z2 w = vload(&W[rr + n * jj]);
    z2 v = vload(&A[rr + n * j]);
    W[rr + n * jj] = w + alpha * v;
  }
}

// ---------------------------------------------------------------- her2k
// A(p:n, p:n) -= V W^H + W V^H, lower tiles of 32x32, V = A(p:n, i0:i0+nbp).
__global__ void __launch_bounds__(256) k_trd_her2k(i64 n, i64 p, i64 i0, int nbp, z2* A_,
                                                   size_t sA, const z2* W_) {
  const int k = blockIdx.y;
  z2* A = A_ + k * sA;
  const z2* W = W_ + (size_t)k * n * NB;
  i64 t = blockIdx.x;
  i64 I = (i64)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= t) ++I;
  while (I * (I + 1) / 2 > t) --I;
  i64 J = t - I * (I + 1) / 2;
  const i64 r0 = p + I * 32, c0 = p + J * 32;
  __shared__ z2 Vr[32][9], Wr[32][9], Vc[32][9], Wc[32][9];
  const int tx = threadIdx.x & 31, g = threadIdx.x >> 5;
  z2 acc[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
  for (int q0 = 0; q0 < nbp; q0 += 8) {
    __syncthreads();
    {
      int q = q0 + g;  // 8 groups load 8 panel columns
      i64 r = r0 + tx, c = c0 + tx;
      z2 z = mk(0.0, 0.0);
      Vr[tx][g] = (q < nbp && r < n) ? A[r + n * (i0 + q)] : z;
      Wr[tx][g] = (q < nbp && r < n) ? W[r + n * q] : z;
      Vc[tx][g] = (q < nbp && c < n) ? A[c + n * (i0 + q)] : z;
      Wc[tx][g] = (q < nbp && c < n) ? W[c + n * q] : z;
    }
    __syncthreads();
    for (int e4 = 0; e4 < 4; ++e4) {
      int cc = g * 4 + e4;
      z2 a = acc[e4];
      for (int q = 0; q < 8; ++q) {
        a = a + mulc(Vr[tx][q], Wc[cc][q]);
        a = a + mulc(Wr[tx][q], Vc[cc][q]);
      }
      acc[e4] = a;
    }
  }
  for (int e4 = 0; e4 < 4; ++e4) {
    i64 r = r0 + tx, c = c0 + g * 4 + e4;
    if (r < n && c < n && r >= c) {
      z2 a = A[r + n * c];
      A[r + n * c] = a - acc[e4];
    }
  }
}

__global__ void k_trd_lastdiag(i64 n, const z2* A_, size_t sA, double* d_) {
  int k = blockIdx.x;
  if (threadIdx.x == 0) d_[(size_t)k * n + n - 1] = A_[k * sA + (n - 1) + n * (n - 1)].x;
}

// ---------------------------------------------------------------- driver
int tridiag(TridiagWork& w, hipStream_t s, double* hemv_ms, long* hemv_launches) {
  const i64 n = w.n;
  const int K = w.K;
  const i64 ntmax = nt_of(n), msmax = seg_of(ntmax), nbmax = nblk_of(n) + 1;
  const bool timed = hemv_ms != nullptr;
  if (timed && w.nev < 2 * n) {
    hipEvent_t* ev = new hipEvent_t[2 * n];
    for (i64 q = 0; q < 2 * n; ++q) FORGE_HIP(hipEventCreate(&ev[q]));
    for (i64 q = 0; q < w.nev; ++q) FORGE_HIP(hipEventDestroy(w.evs[q]));
    delete[] w.evs;
    w.evs = ev;
    w.nev = 2 * n;
  }
  i64 nrec = 0;
  for (i64 i0 = 0; i0 < n - 1; i0 += NB) {
    const int nbp = (int)((n - 1 - i0) < NB ? (n - 1 - i0) : NB);
    for (int jj = 0; jj < nbp; ++jj) {
      const i64 j = i0 + jj;
      hipLaunchKernelGGL(k_trd_col, dim3(cdiv(n - j, 256), K), dim3(256), 0, s, n, j, i0, jj, w.A,
                         w.sA, w.W, w.d, w.e, w.tau, w.npart, w.scal, w.cnt, nbmax);
      FORGE_LAUNCH();
      const i64 m = n - j - 1, nt = nt_of(m), ms = seg_of(nt);
      if (timed) FORGE_HIP(hipEventRecord(w.evs[2 * nrec], s));
      hipLaunchKernelGGL(k_trd_hemv, dim3((unsigned)nt, (unsigned)ms, K), dim3(256), 0, s, n, j,
                         i0, jj, w.A, w.sA, w.W, w.scal, w.P, w.P2, w.tpart, w.t12, w.cnt, ntmax,
                         msmax);
      FORGE_LAUNCH();
      if (timed) {
        FORGE_HIP(hipEventRecord(w.evs[2 * nrec + 1], s));
        ++nrec;
      }
      hipLaunchKernelGGL(k_trd_fin, dim3(cdiv(m, 256), K), dim3(256), 0, s, n, j, i0, jj, w.A,
                         w.sA, w.W, w.scal, w.P, w.P2, w.t12, w.npart, w.cnt, ntmax, msmax, nbmax);
      FORGE_LAUNCH();
    }
    const i64 p = i0 + nbp, mt = n - p, ntt = (mt + 31) / 32;
    hipLaunchKernelGGL(k_trd_her2k, dim3((unsigned)(ntt * (ntt + 1) / 2), K), dim3(256), 0, s, n,
                       p, i0, nbp, w.A, w.sA, w.W);
    FORGE_LAUNCH();
  }
  hipLaunchKernelGGL(k_trd_lastdiag, dim3(K), dim3(64), 0, s, n, w.A, w.sA, w.d);
  FORGE_LAUNCH();
  if (timed && nrec > 0) {
    FORGE_HIP(hipEventSynchronize(w.evs[2 * nrec - 1]));
    for (i64 q = 0; q < nrec; ++q) {
      float ms_ = 0.f;
      FORGE_HIP(hipEventElapsedTime(&ms_, w.evs[2 * q], w.evs[2 * q + 1]));
      *hemv_ms += ms_;
    }
    *hemv_launches += nrec;
  }
  return 0;
}

}  // namespace forge
