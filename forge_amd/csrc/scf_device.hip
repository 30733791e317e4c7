// The host-side steps of one SCF iteration as kernels, so that the whole
// iteration is one launch sequence that a HIP graph can capture:
//
//   k_occupations   SortEnergies_1Spin (HartreeFock.f90:885-925), the Fermi
//                   block (Main.f90:400-412) and the GetFock weights
//                   (HartreeFock.f90:506-555), or Fermi-Dirac occupations at
//                   temperature kT, for the nb x Nk window states in one
//                   block: a bitonic sort of (energy, flat index) in LDS,
//                   the counts and the chemical potential by block
//                   reductions in a fixed order, and the weighted states
//                   compacted by a block scan into the list the window
//                   density reads;
//   k_step_update   OptimalStep's quadratic (HartreeFock.f90:1148-1227) or
//                   linear mixing, the energies of the mixed state and the
//                   trajectory row, from the energy pass's sums;
//   k_density_conv  max |D_mixed - D_in| (Main.f90:564-571);
//   k_mix_dev, k_potential_dev   the mixing and the new potential with the
//                   step read from device memory.
// The formulas are the host loop's (the energy model is shared through
// kernels.hpp); only the order of the floating-point sums differs.
#include "kernels.hpp"

namespace forge {

constexpr int OCC_T = 1024;
constexpr int OCC_W = OCC_T / 64;

__device__ __forceinline__ bool occ_before(double ka, int va, double kb, int vb) {
  return ka < kb || (ka == kb && va < vb);
}

// sum over the block in a fixed order, returned to every thread
__device__ double occ_block_sum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < OCC_W; ++w) s += sh[w];
  return s;
}

__global__ void __launch_bounds__(OCC_T) k_occupations(OccArgs a) {
  __shared__ double key[OCC_MAX];
  __shared__ int val[OCC_MAX];
  __shared__ double dsh[OCC_W];
  __shared__ int cnt[OCC_T];
  const int t = threadIdx.x;
  const int N = a.N;
  int P = 1;
  while (P < N) P <<= 1;
  for (int q = t; q < P; q += OCC_T) {
    key[q] = q < N ? a.band[q] : HUGE_VAL;
    val[q] = q;
  }
  __syncthreads();
  // ascending by (energy, flat index): ties keep the flat order
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int q = t; q < P; q += OCC_T) {
        const int r = q ^ j;
        if (r > q) {
          const bool up = (q & k) == 0;
          const bool r_first = occ_before(key[r], val[r], key[q], val[q]);
          if (r_first == up) {
            const double tk = key[q];
            key[q] = key[r];
            key[r] = tk;
            const int tv = val[q];
            val[q] = val[r];
            val[r] = tv;
          }
        }
      }
      __syncthreads();
    }
  // Fermi block, Main.f90:400-412
  const i64 nF = a.nF;
  const double EF0 = 0.5 * (key[nF - 1] + key[nF]);
  double o = 0.0, pp = 0.0;
  for (int q = t; q < N; q += OCC_T) {
    const double d = key[q] - EF0;
    if (d < -a.energy_tol) o += 1.0;
    else if (fabs(d) < a.energy_tol) { o += 1.0; pp += 1.0; }
  }
  const i64 nOcc = (i64)occ_block_sum(o, dsh);
  const i64 nPart = (i64)occ_block_sum(pp, dsh);
  const double Deg = 1.0 / (double)nPart * (double)(nF + nPart - nOcc);
  double EF = EF0, mu = 0.0;
  if (a.occ_mode == 2) {
    double lo = key[0] - 50 * a.kT - 1.0, hi = key[N - 1] + 50 * a.kT + 1.0;
    for (int it = 0; it < 200; ++it) {
      mu = 0.5 * (lo + hi);
      double sm = 0.0;
      for (int q = t; q < N; q += OCC_T) sm += 1.0 / (1.0 + exp((key[q] - mu) / a.kT));
      sm = occ_block_sum(sm, dsh);
      if (sm > (double)nF) hi = mu;
      else lo = mu;
    }
    EF = mu;
  }
  // weights of a sorted position
  auto weight = [&](int q, double* w, double* wi) {
    *w = 0.0;
    *wi = 0.0;
    if (a.occ_mode == 2) {
      *w = 1.0 / (1.0 + exp((key[q] - mu) / a.kT));
    } else if (q < nOcc) {
      if (q < nOcc - nPart) *w = 1.0;
      else if (a.occ_mode == 0) { *w = (double)(float)Deg; *wi = 8.0; }
      else *w = Deg;
    }
  };
  // compaction in sorted order: thread t owns positions [t c, (t+1) c)
  const int chunk = (N + OCC_T - 1) / OCC_T;
  const int q0 = t * chunk, q1 = min(N, q0 + chunk);
  int mine = 0;
  for (int q = q0; q < q1; ++q) {
    double w, wi;
    weight(q, &w, &wi);
    mine += (w != 0.0 || wi != 0.0) ? 1 : 0;
  }
  cnt[t] = mine;
  __syncthreads();
  for (int off = 1; off < OCC_T; off <<= 1) {  // inclusive scan
    const int v = t >= off ? cnt[t - off] : 0;
    __syncthreads();
    cnt[t] += v;
    __syncthreads();
  }
  int sidx = cnt[t] - mine;
  for (int q = q0; q < q1; ++q) {
    double w, wi;
    weight(q, &w, &wi);
    if (w == 0.0 && wi == 0.0) continue;
    const int src = val[q];
    const int b = src / a.Nk, k = src - b * a.Nk;
    a.idx[sidx] = b + a.nb * k;
    for (int c = 0; c < a.nnc; ++c) {
      const double px = a.ph_cell[(k * a.nnc + c) * 2], py = a.ph_cell[(k * a.nnc + c) * 2 + 1];
      a.wph[((size_t)sidx * a.nnc + c) * 2] = px * w - py * wi;
      a.wph[((size_t)sidx * a.nnc + c) * 2 + 1] = px * wi + py * w;
    }
    ++sidx;
  }
  if (t == OCC_T - 1) *a.S = cnt[t];
  if (t == 0) {
    a.occ[0] = EF;
    a.occ[1] = (double)nOcc;
    a.occ[2] = (double)nPart;
    a.occ[3] = Deg;
  }
}

int device_occupations(const OccArgs& a, hipStream_t s) {
  if (a.N < 2 || a.N > OCC_MAX || a.nF < 1 || a.nF > a.N - 1) {
    set_error("device_occupations: window size or filling out of range");
    return -2;
  }
  hipLaunchKernelGGL(k_occupations, dim3(1), dim3(OCC_T), 0, s, a);
  FORGE_LAUNCH();
  return 0;
}

// band[b Nk + k0 + kb] = evals(nL1 + b) of matrix kb
__global__ void k_band_gather(i64 n, i64 nb, i64 nL1, int K, i64 k0, i64 Nk,
                              const double* __restrict__ evals, double* __restrict__ band) {
  const i64 q = (i64)blockIdx.x * 256 + threadIdx.x;
  if (q >= nb * K) return;
  const i64 b = q % nb, kb = q / nb;
  band[b * Nk + k0 + kb] = evals[kb * n + nL1 + b];
}

int band_gather(i64 n, i64 nb, i64 nL1, int K, i64 k0, i64 Nk, const double* evals, double* band,
                hipStream_t s) {
  hipLaunchKernelGGL(k_band_gather, dim3((unsigned)cdiv(nb * K, 256)), dim3(256), 0, s, n, nb, nL1, K,
                     k0, Nk, evals, band);
  FORGE_LAUNCH();
  return 0;
}

// Psi(:, s) = Cwin(:, idx[s]) for s < *S
__global__ void k_gather_cols_dev(i64 n, i64 Smax, const i64* __restrict__ S,
                                  const int* __restrict__ idx, const z2* __restrict__ Cwin,
                                  z2* __restrict__ Psi) {
  const i64 q = (i64)blockIdx.x * 256 + threadIdx.x;
  if (q >= n * Smax) return;
  const i64 r = q % n, sc = q / n;
  if (sc >= *S) return;
  Psi[q] = Cwin[r + n * (i64)idx[sc]];
}

int gather_cols_dev(i64 n, i64 Smax, const i64* S, const int* idx, const z2* Cwin, z2* Psi,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_gather_cols_dev, dim3((unsigned)cdiv(n * Smax, 256)), dim3(256), 0, s, n, Smax, S,
                     idx, Cwin, Psi);
  FORGE_LAUNCH();
  return 0;
}

__global__ void k_step_update(StepArgs a) {
  if (threadIdx.x != 0) return;
  double pc[13];
  for (int q = 0; q < 5; ++q) pc[q] = a.e5[q];
  for (int q = 0; q < 6; ++q) pc[5 + q] = a.dots[q];
  pc[11] = a.dots[8];
  pc[12] = a.dots[9];
  // st is read and then rewritten by this thread.  Through a plain pointer
  // the compiler loads it on the scalar unit and issues the vector stores
  // below without waiting for those loads (measured: with linear mixing the
  // row's input energies came back as the new ones), so it is read through
  // volatile vector loads, which the stores are ordered after.
  const volatile double* stv = a.st;
  double ein[4] = {stv[0], stv[1], stv[2], stv[3]};
  const double EtotIn = stv[4];
  double e1[4], eh[4], e[4];
  scf_energies_at(pc, 1.0, e1);
  scf_energies_at(pc, 0.5, eh);
  const double step =
      a.mix_mode == 0
          ? scf_oda_step(EtotIn, scf_total_energy(a.alpha, a.alphaH, a.U, e1),
                         scf_total_energy(a.alpha, a.alphaH, a.U, eh), a.step_alt)
          : a.beta;
  scf_energies_at(pc, step, e);
  const double Etot = scf_total_energy(a.alpha, a.alphaH, a.U, e);
  for (int q = 0; q < 4; ++q) a.st[q] = e[q];
  a.st[4] = Etot;
  a.st[5] = step;
  a.row[1] = EtotIn;
  a.row[2] = 2 * ein[0];
  a.row[3] = 2 * ein[1];
  a.row[4] = ein[2];
  a.row[5] = ein[3];
  a.row[7] = step;
  a.row[8] = EtotIn - Etot;
  a.row[9] = a.occ[0];
  a.row[10] = a.occ[1];
  a.row[11] = a.occ[2];
}

int scf_step_update(const StepArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_step_update, dim3(1), dim3(64), 0, s, a);
  FORGE_LAUNCH();
  return 0;
}

__global__ void __launch_bounds__(256) k_density_conv(i64 n, const double* __restrict__ da,
                                                      const double* __restrict__ db,
                                                      const double* __restrict__ st,
                                                      double* __restrict__ row) {
  const double s = st[5];
  double m = 0.0;
  for (i64 i = threadIdx.x; i < n; i += 256) {
    const double dmix = (1.0 - s) * da[i] + s * db[i];
    m = fmax(m, fabs(dmix - da[i]));
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) row[6] = sh[0];
}

int density_convergence(i64 n, const double* da, const double* db, const double* st, double* row,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_density_conv, dim3(1), dim3(256), 0, s, n, da, db, st, row);
  FORGE_LAUNCH();
  return 0;
}

// rho_in <- (1 - s) rho_in + s rho_new, s = st[5]
__global__ void k_mix_dev(size_t tot, z2* __restrict__ a, const z2* __restrict__ b,
                          const double* __restrict__ st) {
  const double s = st[5];
  for (size_t o = (size_t)blockIdx.x * 256 + threadIdx.x; o < tot; o += (size_t)gridDim.x * 256) {
    const z2 x = a[o], y = b[o];
    a[o] = mk(s * y.x + (1.0 - s) * x.x, s * y.y + (1.0 - s) * x.y);
  }
}

int mix_dev(size_t tot, z2* a, const z2* b, const double* st, hipStream_t s) {
  hipLaunchKernelGGL(k_mix_dev, dim3(2048), dim3(256), 0, s, tot, a, b, st);
  FORGE_LAUNCH();
  return 0;
}

// pot = 2 alphaH ((1-s) ya + s yb) + U d, da <- d = (1-s) da + s db
__global__ void k_potential_dev(i64 n, const double* __restrict__ ya, const double* __restrict__ yb,
                                double* __restrict__ da, const double* __restrict__ db,
                                const double* __restrict__ st, double alphaH, double U,
                                double* __restrict__ pot) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double s = st[5];
  const double y = (1.0 - s) * ya[i] + s * yb[i];
  const double d = (1.0 - s) * da[i] + s * db[i];
  pot[i] = 2.0 * alphaH * y + U * d;
  da[i] = d;
}

int potential_dev(i64 n, const double* ya, const double* yb, double* da, const double* db,
                  const double* st, double alphaH, double U, double* pot, hipStream_t s) {
  hipLaunchKernelGGL(k_potential_dev, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, s, n, ya, yb, da, db,
                     st, alphaH, U, pot);
  FORGE_LAUNCH();
  return 0;
}

}  // namespace forge
