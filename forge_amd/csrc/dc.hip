// Batched divide-and-conquer eigensolver for K real symmetric tridiagonal
// matrices (Cuppen's method with Gu-Eisenstat eigenvectors).  It replaces the
// tridiagonal stage of zheevr (LapackRoutines.f90:414-426) in the GPU path.
//
//   tearing   T = diag(T1', T2') + |b| u u^T at every leaf boundary, with
//             T1'(last) -= |b|, T2'(first) -= |b|, u = (e_last, sign(b) e_first)
//   leaves    <= 32 rows: implicit QL with Wilkinson shifts, one wave per leaf,
//             eigenvectors in LDS, then sorted ascending
//   merge     D + rho z z^T with z = (last row of Q1, sign(b) first row of Q2)/sqrt2,
//             rho = 2|b|:
//               1. merge-sort D (ranks by binary search), tol = 8 eps max(|D|,|z|)
//               2. deflate small rho|z_j| and, by a Givens rotation, close pairs
//                  (c = z_q/tau, s = z_p/tau zeroes z_p; the rotation is kept)
//               3. secular roots lam_i in (d_i, d_i+1): bisection in tau with
//                  lam = d_origin + tau (origin = the nearer pole)
//               4. zhat_j^2 = (lam_j - d_j) prod_{i!=j} (lam_i - d_j)/(d_i - d_j) / rho
//               5. u_i = zhat / (d - lam_i) normalised; deflated columns are unit
//                  vectors; the rotations are applied to the rows of W in reverse
//               6. Q = diag(Q1, Q2) W   (two block GEMMs)
//             Only the lowest mout columns are formed at the top merge.
#include <algorithm>
#include <utility>
#include <vector>

#include "kernels.hpp"

namespace forge {

constexpr int LEAF = 32;
// merges of at least this order multiply on the FP64 matrix cores (dgemm,
// gemm.hip, batched over the K matrices, one call per merge half); smaller
// ones on k_dc_gemm, where one launch covers all merges of a level
constexpr int DC_GEMM_MIN = 2048;
constexpr double EPSU = 1.1102230246251565e-16;  // unit roundoff, 2^-53

struct MergeDesc {
  int o, n1, n2, mout;
};

struct DCLayout {
  size_t Qa, Qb, Wm, Ws, Da, Db, perm, Ds, zs, dl, wz, keep, org, taus, zhat, vnorm, colsrc,
      evout, tmp, rp, rq, rc, rs, mk, mrot, mrho, total;
};

static DCLayout dc_layout(i64 n, int K) {
  DCLayout L;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  size_t nn = (size_t)n * n * K, nv = (size_t)n * K;
  L.Qa = take(8 * nn);
  L.Qb = take(8 * nn);
  L.Wm = take(8 * nn);
  L.Ws = take(8 * nn);
  L.Da = take(8 * nv);
  L.Db = take(8 * nv);
  L.perm = take(4 * nv);
  L.Ds = take(8 * nv);
  L.zs = take(8 * nv);
  L.dl = take(8 * nv);
  L.wz = take(8 * nv);
  L.keep = take(4 * nv);
  L.org = take(4 * nv);
  L.taus = take(8 * nv);
  L.zhat = take(8 * nv);
  L.vnorm = take(8 * nv);
  L.colsrc = take(4 * nv);
  L.evout = take(8 * nv);
  L.tmp = take(4 * nv);
  L.rp = take(4 * nv);
  L.rq = take(4 * nv);
  L.rc = take(8 * nv);
  L.rs = take(8 * nv);
  L.mk = take(4 * nv);
  L.mrot = take(4 * nv);
  L.mrho = take(8 * nv);
  L.total = off;
  return L;
}

size_t dc_workspace_bytes(i64 n, int K, i64 mout) {
  (void)mout;
  return dc_layout(n, K).total + 256 * 1024;
}

// ---------------------------------------------------------------- leaves
// One 64-lane block per (leaf, matrix); lanes 0..31 own rows of Z (LDS).
__global__ void __launch_bounds__(64) k_dc_leaf(i64 n, const double* __restrict__ d_,
                                                const double* __restrict__ e_, double* Q_,
                                                double* D_, int* fail) {
  const int k = blockIdx.y;
  const i64 o = (i64)blockIdx.x * LEAF;
  const int ls = (int)((n - o) < LEAF ? (n - o) : LEAF);
  const double* dg = d_ + (size_t)k * n;
  const double* eg = e_ + (size_t)k * n;
  double* Q = Q_ + (size_t)k * n * n;
  __shared__ double Z[LEAF][LEAF + 1];
  __shared__ double d[LEAF], e[LEAF];
  const int t = threadIdx.x;
  if (t < LEAF) {
    for (int c = 0; c < LEAF; ++c) Z[t][c] = (t == c) ? 1.0 : 0.0;
    if (t < ls) {
      double dv = dg[o + t];
      // tearing at the leaf's two boundaries
      if (t == 0 && o > 0) dv -= fabs(eg[o - 1]);
      if (t == ls - 1 && o + ls < n) dv -= fabs(eg[o + ls - 1]);
      d[t] = dv;
      e[t] = (t < ls - 1) ? eg[o + t] : 0.0;
    }
  }
  __syncthreads();
  // implicit QL (all lanes run the scalar recurrence; lanes < ls rotate rows)
  for (int l = 0; l < ls; ++l) {
    int iter = 0;
    for (;;) {
      int m;
      for (m = l; m < ls - 1; ++m) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= EPSU * dd) break;
      }
      if (m == l) break;
      if (++iter > 60) {
        if (t == 0) atomicAdd(fail, 1);
        break;
      }
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      int i;
      bool underflow = false;
      for (i = m - 1; i >= l; --i) {
        double f = s * e[i], b = c * e[i];
        r = hypot(f, g);
        __syncthreads();
        if (t == 0) e[i + 1] = r;
        __syncthreads();
        if (r == 0.0) {
          __syncthreads();
          if (t == 0) {
            d[i + 1] -= p;
            e[m] = 0.0;
          }
          __syncthreads();
          underflow = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        __syncthreads();
        if (t == 0) d[i + 1] = g + p;
        __syncthreads();
        g = c * r - b;
        if (t < ls) {
          double zf = Z[t][i + 1], zi = Z[t][i];
          Z[t][i + 1] = s * zi + c * zf;
          Z[t][i] = c * zi - s * zf;
        }
      }
      if (underflow) continue;
      __syncthreads();
      if (t == 0) {
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
      __syncthreads();
    }
  }
  __syncthreads();
  // selection sort ascending, swapping columns of Z
  for (int a = 0; a < ls - 1; ++a) {
    int mi = a;
    for (int b = a + 1; b < ls; ++b)
      if (d[b] < d[mi]) mi = b;
    __syncthreads();
    if (mi != a) {
      if (t < ls) {
        double tmp = Z[t][a];
        Z[t][a] = Z[t][mi];
        Z[t][mi] = tmp;
      }
      if (t == 0) {
        double tmp = d[a];
        d[a] = d[mi];
        d[mi] = tmp;
      }
    }
    __syncthreads();
  }
  if (t < ls) {
    for (int c = 0; c < ls; ++c) Q[(o + t) + (size_t)n * (o + c)] = Z[t][c];
    D_[(size_t)k * n + o + t] = d[t];
  }
}

// ---------------------------------------------------------------- merge: prep
// One 256-thread block per (merge, matrix).
struct DCPtr {
  double *Qa, *Qb, *Wm, *Ws, *Da, *Db;
  int* perm;
  double *Ds, *zs, *dl, *wz;
  int *keep, *org;
  double *taus, *zhat, *vnorm;
  int* colsrc;
  double* evout;
  int* tmp;
  int *rp, *rq;
  double *rc, *rs;
  int *mk, *mrot;
  double* mrho;
};

__global__ void __launch_bounds__(256) k_dc_prep(i64 n, const MergeDesc* __restrict__ md,
                                                 const double* __restrict__ e_, DCPtr P) {
  const int k = blockIdx.y;
  const MergeDesc M = md[blockIdx.x];
  const int o = M.o, n1 = M.n1, m = M.n1 + M.n2;
  const size_t vo = (size_t)k * n + o;
  const double* Q = P.Qa + (size_t)k * n * n;
  const double* D = P.Da + (size_t)k * n + o;
  double* Ds = P.Ds + vo;
  double* zs = P.zs + vo;
  int* perm = P.perm + vo;
  const double beta = e_[(size_t)k * n + o + n1 - 1];
  const double sgn = beta >= 0.0 ? 1.0 : -1.0;
  const double rho = 2.0 * fabs(beta);
  const double isq2 = 0.70710678118654752440;
  // ranks of the two sorted halves (stable merge)
  for (int t = threadIdx.x; t < m; t += 256) {
    double v, zv;
    int rank;
    if (t < n1) {
      v = D[t];
      zv = Q[(o + n1 - 1) + (size_t)n * (o + t)] * isq2;
      int lo = 0, hi = m - n1;  // count of second-half values < v
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (D[n1 + mid] < v) lo = mid + 1; else hi = mid;
      }
      rank = t + lo;
    } else {
      int u = t - n1;
      v = D[t];
      zv = sgn * Q[(o + n1) + (size_t)n * (o + t)] * isq2;
      int lo = 0, hi = n1;  // count of first-half values <= v
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (D[mid] <= v) lo = mid + 1; else hi = mid;
      }
      rank = u + lo;
    }
    Ds[rank] = v;
    zs[rank] = zv;
    perm[rank] = t;
  }
  __syncthreads();
  __shared__ double sh[256];
  double mx = 0.0, mz = 0.0;
  for (int t = threadIdx.x; t < m; t += 256) {
    mx = fmax(mx, fabs(Ds[t]));
    mz = fmax(mz, fabs(zs[t]));
  }
  sh[threadIdx.x] = fmax(mx, mz);
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  const double tol = 8.0 * EPSU * sh[0];
  __syncthreads();
  sh[threadIdx.x] = mz;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  const double zmax = sh[0];
  if (threadIdx.x != 0) return;
  // sequential deflation scan
  int* keep = P.keep + vo;
  int* rp = P.rp + vo;
  int* rq = P.rq + vo;
  double* rc = P.rc + vo;
  double* rs = P.rs + vo;
  int nk = 0, nr = 0;
  if (!(rho * zmax <= tol)) {
    int prev = -1;
    for (int j = 0; j < m; ++j) {
      if (rho * fabs(zs[j]) <= tol) continue;  // deflated: not kept
      if (prev < 0) {
        prev = j;
        continue;
      }
      double zp = zs[prev], zq = zs[j];
      double tau = hypot(zp, zq);
      double c = zq / tau, s = zp / tau;
      if (fabs((Ds[j] - Ds[prev]) * c * s) <= tol) {
        double dp = Ds[prev], dq = Ds[j];
        Ds[prev] = c * c * dp + s * s * dq;
        Ds[j] = s * s * dp + c * c * dq;
        zs[j] = tau;
        zs[prev] = 0.0;
        rp[nr] = prev;
        rq[nr] = j;
        rc[nr] = c;
        rs[nr] = s;
        ++nr;
      } else {
        keep[nk++] = prev;
      }
      prev = j;
    }
    if (prev >= 0) keep[nk++] = prev;
  }
  double* dl = P.dl + vo;
  double* wz = P.wz + vo;
  for (int i = 0; i < nk; ++i) {
    dl[i] = Ds[keep[i]];
    wz[i] = zs[keep[i]];
  }
  P.mk[vo] = nk;
  P.mrot[vo] = nr;
  P.mrho[vo] = rho;
}

// ---------------------------------------------------------------- merge: secular roots
// Root i of f(t) = 1 + sum_j rho z_j^2 / (delta_j - t), delta_j = d_j - d_org,
// in t = lambda - d_org, org the pole nearer to the root (chosen from the
// sign of f at the interval's midpoint).  f increases from -inf to +inf
// between the poles d_i and d_{i+1} (above d_{nk-1} for the last root).
// Iteration: the two-pole rational model of LAPACK dlaed4's fixed-weight
// scheme (the poles left of the root summed into psi, the rest into phi;
// each part replaced by c + s / (pole - t) matching its value and slope at
// the iterate), whose root is taken when it falls inside the bracket [a, b]
// kept from the signs of f; otherwise (and after 20 steps) a bisection step.
// Stops when |f| is below its rounding bound 8 eps (1 + sum_j |term_j|), the
// step is below 4 eps |t| or the bracket has collapsed: about 5 passes over
// the poles on random and clustered spectra where bisection needs about 55.
// One wave per root (4 roots per 256-thread block): the lanes split every
// sum over the poles, a butterfly reduces it and lane 0's value is
// broadcast, so all lanes take the same branches.
__device__ inline double wave_total(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return __shfl(v, 0, 64);
}

__global__ void __launch_bounds__(256) k_dc_secular(i64 n, const MergeDesc* __restrict__ md,
                                                    DCPtr P) {
  const int k = blockIdx.z;
  const MergeDesc M = md[blockIdx.y];
  const size_t vo = (size_t)k * n + M.o;
  const int nk = P.mk[vo];
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= nk) return;  // wave-uniform
  const double* dl = P.dl + vo;
  const double* wz = P.wz + vo;
  const double rho = P.mrho[vo];
  const bool last = i == nk - 1;
  double lo = dl[i], hi;
  if (!last) {
    hi = dl[i + 1];
  } else {
    double s = 0.0;
    for (int j = lane; j < nk; j += 64) s += wz[j] * wz[j];
    hi = dl[nk - 1] + rho * wave_total(s);
  }
  double half = 0.5 * (hi - lo);
  int org;
  double a, b;
  bool left = last;
  if (!left) {
    double f = 0.0;
    for (int j = lane; j < nk; j += 64) f += rho * wz[j] * wz[j] / ((dl[j] - lo) - half);
    left = 1.0 + wave_total(f) > 0.0;
  }
  if (left) {
    org = i;
    a = 0.0;
    b = last ? (hi - lo) : half;
  } else {
    org = i + 1;
    a = -half;
    b = 0.0;
  }
  const double dorg = dl[org];
  // the poles bracketing the root, in t
  const double pL = dl[i] - dorg, pR = last ? 0.0 : dl[i + 1] - dorg;
  constexpr double eps = 1.1102230246251565e-16;
  double t = 0.5 * (a + b);
  for (int it = 0; it < 120; ++it) {
    double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0, absum = 0.0;
    for (int j = lane; j < nk; j += 64) {
      const double del = (dl[j] - dorg) - t;
      const double term = rho * wz[j] * wz[j] / del;
      const double der = term / del;
      if (j <= i) { psi += term; dpsi += der; }
      else { phi += term; dphi += der; }
      absum += fabs(term);
    }
    psi = wave_total(psi); dpsi = wave_total(dpsi);
    phi = wave_total(phi); dphi = wave_total(dphi);
    absum = wave_total(absum);
    const double f = 1.0 + psi + phi;
    if (fabs(f) <= 8.0 * eps * (1.0 + absum)) break;
    if (f > 0.0) b = t; else a = t;
    const double mid = 0.5 * (a + b);
    if (mid == a || mid == b) break;
    double tn = mid;
    if (it < 20) {
      const double D1 = pL - t;
      double eta = 0.0;
      bool ok = false;
      if (last) {
        const double c = f - dpsi * D1;
        if (c > 0.0) { eta = D1 + dpsi * D1 * D1 / c; ok = true; }
      } else {
        const double D2 = pR - t;
        const double c = f - dpsi * D1 - dphi * D2;
        const double A = c * (D1 + D2) + dpsi * D1 * D1 + dphi * D2 * D2;
        const double Bc = D1 * D2 * f;
        if (c == 0.0) {
          if (A != 0.0) { eta = Bc / A; ok = true; }
        } else {
          const double disc = sqrt(fabs(A * A - 4.0 * Bc * c));
          if (A <= 0.0) eta = (A - disc) / (2.0 * c);
          else eta = 2.0 * Bc / (A + disc);
          ok = true;
        }
      }
      const double tr = t + eta;
      if (ok && tr > a && tr < b) {
        tn = tr;
        if (fabs(eta) <= 4.0 * eps * fabs(t)) { t = tn; break; }
      }
    }
    t = tn;
  }
  if (t == 0.0) t = (b != 0.0) ? b : a;
  if (lane == 0) {
    P.org[vo + i] = org;
    P.taus[vo + i] = t;
  }
}

// zhat (Gu-Eisenstat) and then the column norms of the secular eigenvectors.
__global__ void __launch_bounds__(256) k_dc_zhat(i64 n, const MergeDesc* __restrict__ md, DCPtr P) {
  const int k = blockIdx.z;
  const MergeDesc M = md[blockIdx.y];
  const size_t vo = (size_t)k * n + M.o;
  const int nk = P.mk[vo];
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= nk) return;
  const double* dl = P.dl + vo;
  const int* org = P.org + vo;
  const double* taus = P.taus + vo;
  const double dj = dl[j];
  double prod = 1.0;
  for (int i = 0; i < nk; ++i) {
    double lam_m_dj = taus[i] - (dj - dl[org[i]]);  // lam_i - d_j
    if (i == j) prod *= lam_m_dj;
    else prod *= lam_m_dj / (dl[i] - dj);
  }
  double v = prod / P.mrho[vo];
  P.zhat[vo + j] = copysign(sqrt(v > 0.0 ? v : 0.0), P.wz[vo + j]);
}

__global__ void __launch_bounds__(256) k_dc_vnorm(i64 n, const MergeDesc* __restrict__ md, DCPtr P) {
  const int k = blockIdx.z;
  const MergeDesc M = md[blockIdx.y];
  const size_t vo = (size_t)k * n + M.o;
  const int nk = P.mk[vo];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nk) return;
  const double* dl = P.dl + vo;
  const double* zh = P.zhat + vo;
  const double dorg = dl[P.org[vo + i]], t = P.taus[vo + i];
  double s = 0.0;
  for (int j = 0; j < nk; ++j) {
    double u = zh[j] / ((dl[j] - dorg) - t);
    s += u * u;
  }
  P.vnorm[vo + i] = sqrt(s);
}

// Column order of the merged problem: the deflated values (Ds at the
// non-kept positions, nearly sorted -- Givens deflation moves them by at most
// tol) are insertion-sorted, then merged with the roots lam_i (increasing by
// interlacing).  colsrc[c] = root index i (>= 0) or -(p+1) for position p.
// One thread per merge: O(m) for nearly sorted input.
__global__ void k_dc_order(i64 n, const MergeDesc* __restrict__ md, DCPtr P) {
  const int k = blockIdx.y;
  const MergeDesc M = md[blockIdx.x];
  if (threadIdx.x != 0) return;
  const int m = M.n1 + M.n2;
  const size_t vo = (size_t)k * n + M.o;
  const int nk = P.mk[vo];
  const int* keep = P.keep + vo;
  const double* Ds = P.Ds + vo;
  const double* dl = P.dl + vo;
  const int* org = P.org + vo;
  const double* taus = P.taus + vo;
  int* dpos = P.tmp + vo;
  int nd = 0, kp = 0;
  for (int p = 0; p < m; ++p) {
    if (kp < nk && keep[kp] == p) { ++kp; continue; }
    // insertion into the sorted deflated list
    int q = nd++;
    double v = Ds[p];
    while (q > 0 && Ds[dpos[q - 1]] > v) { dpos[q] = dpos[q - 1]; --q; }
    dpos[q] = p;
  }
  int a = 0, b = 0;
  for (int c = 0; c < m; ++c) {
    double lam = b < nk ? dl[org[b]] + taus[b] : 0.0;
    bool take_defl = a < nd && (b >= nk || Ds[dpos[a]] <= lam);
    if (take_defl) {
      P.colsrc[vo + c] = -(dpos[a] + 1);
      P.evout[vo + c] = Ds[dpos[a]];
      ++a;
    } else {
      P.colsrc[vo + c] = b;
      P.evout[vo + c] = lam;
      ++b;
    }
  }
}

// W column c (sorted basis in Ws, then rotated, then scattered by perm into Wm).
// One 64-thread block per (column, merge, matrix).
__global__ void __launch_bounds__(64) k_dc_buildW(i64 n, const MergeDesc* __restrict__ md, DCPtr P) {
  const int k = blockIdx.z;
  const MergeDesc M = md[blockIdx.y];
  const int c = blockIdx.x;
  if (c >= M.mout) return;
  const int o = M.o, m = M.n1 + M.n2;
  const size_t vo = (size_t)k * n + o;
  const int nk = P.mk[vo];
  double* x = P.Ws + (size_t)k * n * n + o + (size_t)n * (o + c);
  double* w = P.Wm + (size_t)k * n * n + o + (size_t)n * (o + c);
  const int src = P.colsrc[vo + c];
  const int* keep = P.keep + vo;
  for (int r = threadIdx.x; r < m; r += 64) x[r] = 0.0;
  __syncthreads();
  if (src >= 0) {
    const double* dl = P.dl + vo;
    const double* zh = P.zhat + vo;
    const double dorg = dl[P.org[vo + src]], t = P.taus[vo + src];
    const double inv = 1.0 / P.vnorm[vo + src];
    for (int j = threadIdx.x; j < nk; j += 64) x[keep[j]] = zh[j] / ((dl[j] - dorg) - t) * inv;
  } else {
    if (threadIdx.x == 0) x[-src - 1] = 1.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nr = P.mrot[vo];
    const int* rp = P.rp + vo;
    const int* rq = P.rq + vo;
    const double* rc = P.rc + vo;
    const double* rs = P.rs + vo;
    for (int q = nr - 1; q >= 0; --q) {
      double a = x[rp[q]], b = x[rq[q]];
      x[rp[q]] = rc[q] * a + rs[q] * b;
      x[rq[q]] = -rs[q] * a + rc[q] * b;
    }
  }
  __syncthreads();
  const int* perm = P.perm + vo;
  for (int r = threadIdx.x; r < m; r += 64) w[perm[r]] = x[r];
}

// Q_out(o + r, o + c) = sum_t Qin(o + r, o + t) W(o + t, o + c) over the block
// half that contains r.  64x64 output tiles, LDS-staged 16-deep panels.
__global__ void __launch_bounds__(256) k_dc_gemm(i64 n, const MergeDesc* __restrict__ md,
                                                 const double* __restrict__ Qin_,
                                                 const double* __restrict__ W_, double* Qout_,
                                                 i64 ldo) {
  const int k = blockIdx.z;
  const MergeDesc M = md[blockIdx.y >> 1];
  const int half = blockIdx.y & 1;
  const int o = M.o, m = M.n1 + M.n2;
  const int rb = half ? M.n1 : 0, rn = half ? M.n2 : M.n1;  // rows/inner range of this half
  const int tilesR = (rn + 63) / 64, tilesC = (M.mout + 63) / 64;
  if ((int)blockIdx.x >= tilesR * tilesC) return;
  const int tr = blockIdx.x % tilesR, tc = blockIdx.x / tilesR;
  const double* Qin = Qin_ + (size_t)k * n * n;
  const double* W = W_ + (size_t)k * n * n;
  double* Qout = Qout_ + (size_t)k * ldo * n;
  __shared__ double As[16][64 + 1], Bs[16][64 + 1];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;  // 16 x 16 threads, 4x4 each
  double acc[4][4] = {};
  for (int k0 = 0; k0 < rn; k0 += 16) {
    __syncthreads();
    for (int q = threadIdx.x; q < 16 * 64; q += 256) {
      int kk = q >> 6, rr = q & 63;
      int row = tr * 64 + rr, kin = k0 + kk;
      As[kk][rr] = (row < rn && kin < rn) ? Qin[(o + rb + row) + (size_t)n * (o + rb + kin)] : 0.0;
      int col = tc * 64 + rr;
      Bs[kk][rr] = (col < M.mout && kin < rn) ? W[(o + rb + kin) + (size_t)n * (o + col)] : 0.0;
    }
    __syncthreads();
    for (int kk = 0; kk < 16; ++kk) {
      double a[4], b[4];
      for (int u = 0; u < 4; ++u) a[u] = As[kk][tx + 16 * u];
      for (int v = 0; v < 4; ++v) b[v] = Bs[kk][ty + 16 * v];
      for (int u = 0; u < 4; ++u)
        for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * b[v];
    }
  }
  for (int u = 0; u < 4; ++u)
    for (int v = 0; v < 4; ++v) {
      int row = tr * 64 + tx + 16 * u, col = tc * 64 + ty + 16 * v;
      if (row < rn && col < M.mout) Qout[(o + rb + row) + (size_t)ldo * (o + col)] = acc[u][v];
    }
  (void)m;
}

__global__ void k_dc_evcopy(i64 n, const MergeDesc* __restrict__ md, DCPtr P, double* Dout) {
  const int k = blockIdx.z;
  const MergeDesc M = md[blockIdx.y];
  const size_t vo = (size_t)k * n + M.o;
  int c = blockIdx.x * 256 + threadIdx.x;
  if (c < M.mout) Dout[vo + c] = P.evout[vo + c];
}

// copy an unmerged block (o, m) from (Da, Qa) to (Db, Qb)
__global__ void k_dc_carry(i64 n, int o, int m, const double* Qa, const double* Da, double* Qb,
                           double* Db) {
  const int k = blockIdx.y;
  i64 q = (i64)blockIdx.x * 256 + threadIdx.x;
  if (q >= (i64)m * m) return;
  int r = (int)(q % m), c = (int)(q / m);
  size_t off = (size_t)k * n * n + (o + r) + (size_t)n * (o + c);
  Qb[off] = Qa[off];
  if (c == 0) Db[(size_t)k * n + o + r] = Da[(size_t)k * n + o + r];
}

__global__ void k_dc_copyout(i64 n, i64 mout, const double* Qa, const double* Da, double* Qo,
                             double* Do) {
  const int k = blockIdx.y;
  i64 q = (i64)blockIdx.x * 256 + threadIdx.x;
  if (q >= n * mout) return;
  Qo[(size_t)k * n * mout + q] = Qa[(size_t)k * n * n + q];
  if (q < mout) Do[(size_t)k * n + q] = Da[(size_t)k * n + q];
}

__global__ void k_dc_fill(size_t count, double* __restrict__ p, int* __restrict__ flag) {
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < count; q += (size_t)gridDim.x * 256)
    p[q] = 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 0;
}

const int* dc_fail_ptr(void* work, i64 n, int K) {
  return (const int*)((char*)work + dc_layout(n, K).total + 200 * 1024);
}

int dc_solve(const double* d, const double* e, i64 n, int K, i64 mout, double* evals, double* Qout,
             void* work, hipStream_t s, bool deferred) {
  DCLayout L = dc_layout(n, K);
  char* b = (char*)work;
  DCPtr P;
  P.Qa = (double*)(b + L.Qa); P.Qb = (double*)(b + L.Qb);
  P.Wm = (double*)(b + L.Wm); P.Ws = (double*)(b + L.Ws);
  P.Da = (double*)(b + L.Da); P.Db = (double*)(b + L.Db);
  P.perm = (int*)(b + L.perm); P.Ds = (double*)(b + L.Ds); P.zs = (double*)(b + L.zs);
  P.dl = (double*)(b + L.dl); P.wz = (double*)(b + L.wz); P.keep = (int*)(b + L.keep);
  P.org = (int*)(b + L.org); P.taus = (double*)(b + L.taus); P.zhat = (double*)(b + L.zhat);
  P.vnorm = (double*)(b + L.vnorm); P.colsrc = (int*)(b + L.colsrc);
  P.evout = (double*)(b + L.evout); P.tmp = (int*)(b + L.tmp); P.rp = (int*)(b + L.rp); P.rq = (int*)(b + L.rq);
  P.rc = (double*)(b + L.rc); P.rs = (double*)(b + L.rs); P.mk = (int*)(b + L.mk);
  P.mrot = (int*)(b + L.mrot); P.mrho = (double*)(b + L.mrho);
  MergeDesc* dmd = (MergeDesc*)(b + L.total);
  int* dfail = (int*)(b + L.total + 200 * 1024);
  if (deferred) {
    // kernels, not memset nodes, inside a captured graph
    hipLaunchKernelGGL(k_dc_fill, dim3(2048), dim3(256), 0, s, (size_t)n * n * K, P.Qa, dfail);
    FORGE_LAUNCH();
  } else {
    FORGE_HIP(hipMemsetAsync(dfail, 0, sizeof(int), s));
    FORGE_HIP(hipMemsetAsync(P.Qa, 0, sizeof(double) * n * n * K, s));
  }
  // leaves
  const int nleaf = (int)((n + LEAF - 1) / LEAF);
  hipLaunchKernelGGL(k_dc_leaf, dim3(nleaf, K), dim3(64), 0, s, n, d, e, P.Qa, P.Da, dfail);
  FORGE_LAUNCH();
  // merge schedule
  struct Blk { int o, m; };
  std::vector<Blk> blocks;
  for (int q = 0; q < nleaf; ++q) {
    int o = q * LEAF;
    blocks.push_back({o, (int)((n - o) < LEAF ? (n - o) : LEAF)});
  }
  std::vector<MergeDesc> all;
  std::vector<std::pair<size_t, size_t>> levels;  // (offset into all, count)
  std::vector<std::vector<Blk>> carries;
  while (blocks.size() > 1) {
    std::vector<Blk> nb;
    size_t off = all.size();
    bool top = blocks.size() == 2;
    std::vector<Blk> carry;
    for (size_t q = 0; q + 1 < blocks.size(); q += 2) {
      MergeDesc M;
      M.o = blocks[q].o;
      M.n1 = blocks[q].m;
      M.n2 = blocks[q + 1].m;
      M.mout = top ? (int)mout : M.n1 + M.n2;
      all.push_back(M);
      nb.push_back({M.o, M.n1 + M.n2});
    }
    if (blocks.size() % 2) {
      nb.push_back(blocks.back());
      carry.push_back(blocks.back());
    }
    levels.push_back({off, all.size() - off});
    carries.push_back(carry);
    blocks.swap(nb);
  }
  if (!all.empty()) {
    if (all.size() * sizeof(MergeDesc) > 200 * 1024) { set_error("dc: too many merges"); return -2; }
    // deferred calls reuse the descriptors a non-deferred call on the same
    // workspace and shape uploaded (they depend on n and mout only): a copy
    // from this host vector could not be replayed from a graph
    if (!deferred)
      FORGE_HIP(hipMemcpyAsync(dmd, all.data(), all.size() * sizeof(MergeDesc), hipMemcpyHostToDevice, s));
  }
  int maxm = 0;
  for (auto& M : all) maxm = std::max(maxm, M.n1 + M.n2);
  for (size_t lv = 0; lv < levels.size(); ++lv) {
    const MergeDesc* md = dmd + levels[lv].first;
    const unsigned nm = (unsigned)levels[lv].second;
    int lm = 0, lmout = 0, lhalf = 0;
    for (size_t q = 0; q < nm; ++q) {
      const MergeDesc& M = all[levels[lv].first + q];
      lm = std::max(lm, M.n1 + M.n2);
      lmout = std::max(lmout, M.mout);
      lhalf = std::max(lhalf, std::max(M.n1, M.n2));
    }
    const bool top = lv + 1 == levels.size();
    hipLaunchKernelGGL(k_dc_prep, dim3(nm, K), dim3(256), 0, s, n, md, e, P);
    FORGE_LAUNCH();
    dim3 gr(cdiv(lm, 256), nm, K);
    hipLaunchKernelGGL(k_dc_secular, dim3(cdiv(lm, 4), nm, K), dim3(256), 0, s, n, md, P);
    FORGE_LAUNCH();
    hipLaunchKernelGGL(k_dc_zhat, gr, dim3(256), 0, s, n, md, P);
    FORGE_LAUNCH();
    hipLaunchKernelGGL(k_dc_vnorm, gr, dim3(256), 0, s, n, md, P);
    FORGE_LAUNCH();
    hipLaunchKernelGGL(k_dc_order, dim3(nm, K), dim3(64), 0, s, n, md, P);
    FORGE_LAUNCH();
    hipLaunchKernelGGL(k_dc_buildW, dim3(lmout, nm, K), dim3(64), 0, s, n, md, P);
    FORGE_LAUNCH();
    double* dst = top ? Qout : P.Qb;
    i64 ldo = n;
    // Q_out for the top merge is n x mout with ld n (caller buffer K x n x mout)
    if (top) {
      // the top merge is the bulk of the work: its two half products go to
      // the FP64 matrix cores (dgemm), all K matrices in one batched call
      // each; caller layout: stride n * mout per matrix
      if (nm == 1) {
        const MergeDesc& M = all[levels[lv].first];
        for (int half = 0; half < 2; ++half) {
          const int rb = half ? M.n1 : 0, rn = half ? M.n2 : M.n1;
          if (rn == 0 || M.mout == 0) continue;
          const size_t a0 = (size_t)(M.o + rb) + (size_t)n * (M.o + rb);
          const size_t b0 = (size_t)(M.o + rb) + (size_t)n * M.o;
          const size_t c0 = (size_t)(M.o + rb) + (size_t)ldo * M.o;
          const int r = dgemm(Op::N, Op::N, rn, M.mout, rn, 1.0, P.Qa + a0, n, (size_t)n * n,
                              P.Wm + b0, n, (size_t)n * n, 0.0, Qout + c0, ldo, (size_t)n * mout, K, s);
          if (r) return r;
        }
      } else {
        unsigned tiles = (unsigned)(((lhalf + 63) / 64) * ((lmout + 63) / 64));
        for (int k = 0; k < K; ++k) {
          hipLaunchKernelGGL(k_dc_gemm, dim3(tiles, 2 * nm, 1), dim3(256), 0, s, n, md,
                             P.Qa + (size_t)k * n * n, P.Wm + (size_t)k * n * n,
                             Qout + (size_t)k * n * mout, ldo);
          FORGE_LAUNCH();
        }
      }
      hipLaunchKernelGGL(k_dc_evcopy, dim3(cdiv(lmout, 256), nm, K), dim3(256), 0, s, n, md, P,
                         evals);
      FORGE_LAUNCH();
    } else {
      if (lm >= DC_GEMM_MIN) {
        // large lower merges: their half products on the matrix cores as well,
        // one call per merge and half, batched over the K matrices
        const size_t snn = (size_t)n * n;
        for (size_t q = 0; q < nm; ++q) {
          const MergeDesc& M = all[levels[lv].first + q];
          for (int half = 0; half < 2; ++half) {
            const int rb = half ? M.n1 : 0, rn = half ? M.n2 : M.n1;
            if (rn == 0 || M.mout == 0) continue;
            const size_t a0 = (size_t)(M.o + rb) + (size_t)n * (M.o + rb);
            const size_t b0 = (size_t)(M.o + rb) + (size_t)n * M.o;
            const int r = dgemm(Op::N, Op::N, rn, M.mout, rn, 1.0, P.Qa + a0, n, snn, P.Wm + b0, n,
                                snn, 0.0, dst + b0, ldo, snn, K, s);
            if (r) return r;
          }
        }
      } else {
        unsigned tiles = (unsigned)(((lhalf + 63) / 64) * ((lmout + 63) / 64));
        hipLaunchKernelGGL(k_dc_gemm, dim3(tiles, 2 * nm, K), dim3(256), 0, s, n, md, P.Qa, P.Wm,
                           dst, ldo);
        FORGE_LAUNCH();
      }
      hipLaunchKernelGGL(k_dc_evcopy, dim3(cdiv(lmout, 256), nm, K), dim3(256), 0, s, n, md, P,
                         P.Db);
      FORGE_LAUNCH();
      for (auto& c : carries[lv]) {
        hipLaunchKernelGGL(k_dc_carry, dim3(cdiv((i64)c.m * c.m, 256), K), dim3(256), 0, s, n,
                           c.o, c.m, P.Qa, P.Da, P.Qb, P.Db);
        FORGE_LAUNCH();
      }
      std::swap(P.Qa, P.Qb);
      std::swap(P.Da, P.Db);
    }
  }
  if (levels.empty()) {
    hipLaunchKernelGGL(k_dc_copyout, dim3(cdiv(n * mout, 256), K), dim3(256), 0, s, n, mout, P.Qa,
                       P.Da, Qout, evals);
    FORGE_LAUNCH();
  }
  if (deferred) return 0;
  int hfail = 0;
  FORGE_HIP(hipMemcpyAsync(&hfail, dfail, sizeof(int), hipMemcpyDeviceToHost, s));
  FORGE_HIP(hipStreamSynchronize(s));
  return hfail;  // > 0: leaves whose QL iteration did not converge
}

}  // namespace forge
