// Density-matrix products of Solve / GetFock (HartreeFock.f90:456-475, 487-560).
//
// Bulk (bands 1..nLower-1 of every k): the reference accumulates nLower-1
// rank-1 zgemm updates per k, T(i,j) = sum_b conj(psi_b(i)) psi_b(j), and adds
// T e^{-2 pi i (i1 n1 + i2 (n1 + n2)) / numk} to every stored cell.  Here
// R_k = C_occ C_occ^H is one rank-nocc Hermitian product per k on the FP64
// matrix cores (gemm.hip:herk_cells), whose epilogue adds the phased
// T = conj(R_k) to every cell directly, both triangles from the lower tiles,
// so R_k never goes through memory.
//
// Window (sorted states s of the 20-band window): the reference adds
// w_s e^{-i k_s.T_c} conj(psi_s) psi_s^T per state and cell with w_s = 1 for
// occupied states and zdeg for the degenerate shell.  k_window_density forms
// all cells for all states in one LDS-tiled pass and writes
// rho_new = (bulk + window) / Nk (Main.f90:510-511).
#include "kernels.hpp"

namespace forge {

// rho (nnc cells) receives sum_k ph_c(k) conj(R_k); accumulate = false
// overwrites at the first k.  The product stays FP64 in every phase: rounding
// of R enters the energies at first order (E_kin and the Hartree terms are
// linear in rho).
int bulk_density(i64 n, int K, const z2* C, size_t ldC, size_t sC, i64 nocc, int nnc,
                 const double* phase_cells, z2* rho, bool accumulate, hipStream_t s) {
  const size_t nn = (size_t)n * n;
  for (int k = 0; k < K; ++k) {
    const bool first = k == 0 && !accumulate;
    if (nocc > 0) {
      const int r = herk_cells(n, nocc, C + k * sC, (i64)ldC, nnc, phase_cells + (size_t)k * nnc * 2,
                               rho, first, s);
      if (r) return r;
    } else if (first) {
      FORGE_HIP(hipMemsetAsync(rho, 0, sizeof(z2) * nnc * nn, s));
    }
  }
  return 0;
}

// rho_out_c(i,j) = scale * (bulk_c(i,j) + sum_s wph[s][c] conj(Psi(i,s)) Psi(j,s)).
// Psi: n x S (ld n); wph: [S][nnc][2] complex weights (occupation x cell phase).
// 32x32 output tiles, states staged 16 at a time.  bulk and out may alias.
// Sdev non-null: the state count is read from device memory (it is computed
// on the device in the graph-captured loop).
template <int NNC>
__global__ void __launch_bounds__(256) k_window_density(i64 n, i64 S, const i64* Sdev,
                                                        const z2* __restrict__ Psi,
                                                        const double* __restrict__ wph,
                                                        const z2* bulk, z2* out, double scale) {
  if (Sdev) S = *Sdev;
  const size_t nn = (size_t)n * n;
  const i64 I = blockIdx.y, J = blockIdx.x;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  __shared__ z2 Pi[16][33], Pj[16][33];
  constexpr int nnc = NNC;
  __shared__ z2 W[16][NNC];
  z2 acc[4][NNC];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int c = 0; c < NNC; ++c) acc[q][c] = mk(0.0, 0.0);
  for (i64 s0 = 0; s0 < S; s0 += 16) {
    __syncthreads();
    for (int q = threadIdx.x; q < 16 * 32; q += 256) {
      int ss = q >> 5, rr = q & 31;
      i64 sidx = s0 + ss;
      i64 ri = I * 32 + rr, rj = J * 32 + rr;
      Pi[ss][rr] = (sidx < S && ri < n) ? conjz(Psi[ri + n * sidx]) : mk(0.0, 0.0);
      Pj[ss][rr] = (sidx < S && rj < n) ? Psi[rj + n * sidx] : mk(0.0, 0.0);
    }
    if (threadIdx.x < 16 * nnc) {
      int ss = threadIdx.x / nnc, c = threadIdx.x % nnc;
      i64 sidx = s0 + ss;
      W[ss][c] = sidx < S ? mk(wph[(sidx * nnc + c) * 2], wph[(sidx * nnc + c) * 2 + 1])
                          : mk(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll 4
    for (int ss = 0; ss < 16; ++ss) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        z2 pr = Pi[ss][tx] * Pj[ss][ty * 4 + q];
#pragma unroll
        for (int c = 0; c < NNC; ++c) acc[q][c] = acc[q][c] + W[ss][c] * pr;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    i64 i = I * 32 + tx, j = J * 32 + ty * 4 + q;
    if (i >= n || j >= n) continue;
    size_t o = i + n * j;
#pragma unroll
    for (int c = 0; c < NNC; ++c) {
      z2 b = bulk ? bulk[c * nn + o] : mk(0.0, 0.0);
      out[c * nn + o] = scale * (b + acc[q][c]);
    }
  }
}

int window_density(i64 n, int nnc, i64 S, const z2* Psi, const double* wph, z2* rho,
                   double scale, hipStream_t s, const i64* Sdev) {
  dim3 grid(cdiv(n, 32), cdiv(n, 32));
  if (nnc != 4) { set_error("window_density: 4 stored cells (numI = 1) expected"); return -2; }
  hipLaunchKernelGGL(k_window_density<4>, grid, dim3(256), 0, s, n, S, Sdev, Psi, wph, rho, rho,
                     scale);
  FORGE_LAUNCH();
  return 0;
}

}  // namespace forge
