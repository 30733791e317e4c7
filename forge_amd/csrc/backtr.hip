// Back-transformation C = Q Z of the tridiagonal eigenvectors Z (real, n x m)
// to eigenvectors of the Hermitian matrices, Q = H_0 ... H_{n-2} from
// tridiag.hip (zunmtr('L','L','N') semantics; zheevr applies it in
// LapackRoutines.f90:414-426).  Reflectors are applied in blocks of BT in the
// compact WY form  H_i0 ... H_{i0+b-1} = I - V T V^H  (T upper triangular,
// LAPACK zlarft 'F','C'), last block first:
//     Y = V^H C,  Y <- T Y,  C <- C - V Y          (rocBLAS zgemm, batched over k)
// The clean V (unit diagonal, zeros above) and T are built by small kernels.
#include <rocblas/rocblas.h>

#include "kernels.hpp"

namespace forge {

constexpr int BT = 128;

size_t backtr_workspace_bytes(i64 n, int K, i64 m) {
  // V (n x BT), G (BT x BT), T (BT x BT), Y (BT x m), Y2 (BT x m) per matrix
  size_t per = (size_t)n * BT + 2 * BT * BT + 2 * (size_t)BT * m;
  return sizeof(z2) * per * K + 4096;
}

__global__ void k_bt_promote(size_t nm, const double* __restrict__ Z, size_t ldZ, size_t sZ,
                             i64 n, i64 m, z2* __restrict__ C, size_t ldC, size_t sC) {
  int k = blockIdx.y;
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < nm; q += (size_t)gridDim.x * 256) {
    i64 r = (i64)(q % n), c = (i64)(q / n);
    C[k * sC + r + ldC * c] = mk(Z[k * sZ + r + ldZ * c], 0.0);
  }
}

// V(r', q) for rows r = i0+1+r' of reflector j = i0+q: 1 at r = j+1, A(r, j) below.
__global__ void k_bt_vclean(i64 n, i64 i0, int nb, const z2* __restrict__ A, size_t sA, i64 lda,
                            z2* __restrict__ V, size_t sV) {
  int k = blockIdx.y;
  i64 mt = n - i0 - 1;
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < (size_t)mt * nb;
       q += (size_t)gridDim.x * 256) {
    i64 rr = (i64)(q % mt), c = (i64)(q / mt);
    i64 r = i0 + 1 + rr, j = i0 + c;
    z2 v;
    if (r == j + 1) v = mk(1.0, 0.0);
    else if (r > j + 1) v = A[k * sA + r + lda * j];
    else v = mk(0.0, 0.0);
    V[k * sV + rr + (size_t)n * c] = v;
  }
}

// T (nb x nb, upper) from G = V^H V and tau, zlarft forward/columnwise:
//   T(i,i) = tau_i,  T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i).
__global__ void k_bt_tbuild(int nb, i64 i0, i64 n, const z2* __restrict__ G, const z2* __restrict__ tau,
                            z2* __restrict__ T) {
  int k = blockIdx.x;
  const z2* g = G + (size_t)k * BT * BT;
  z2* t = T + (size_t)k * BT * BT;
  const z2* tk = tau + (size_t)k * n + i0;
  __shared__ z2 col[BT];
  for (int q = threadIdx.x; q < BT * BT; q += blockDim.x) t[q] = mk(0.0, 0.0);
  __syncthreads();
  for (int i = 0; i < nb; ++i) {
    // col(0:i) = -tau_i G(0:i, i)
    for (int r = threadIdx.x; r < i; r += blockDim.x) col[r] = (-1.0) * (tk[i] * g[r + BT * i]);
    __syncthreads();
    // T(0:i, i) = T(0:i,0:i) col
    for (int r = threadIdx.x; r < i; r += blockDim.x) {
      z2 s = mk(0.0, 0.0);
      for (int c = r; c < i; ++c) s = s + t[r + BT * c] * col[c];
      t[r + BT * i] = s;
    }
    if (threadIdx.x == 0) t[i + BT * i] = tk[i];
    __syncthreads();
  }
}

#define RB(call)                                                              \
  do {                                                                        \
    rocblas_status _s = (call);                                               \
    if (_s != rocblas_status_success) {                                       \
      set_error(std::string("rocblas: ") + rocblas_status_to_string(_s) + " (" + #call + ")"); \
      return -1;                                                              \
    }                                                                         \
  } while (0)

rocblas_handle blas_handle(hipStream_t s);

int backtransform(const TridiagWork& t, const double* Z, size_t ldZ, size_t sZ, i64 m, z2* C,
                  size_t ldC, size_t sC, void* work, hipStream_t s) {
  const i64 n = t.n;
  const int K = t.K;
  z2* V = (z2*)work;
  z2* G = V + (size_t)n * BT * K;
  z2* T = G + (size_t)BT * BT * K;
  z2* Y = T + (size_t)BT * BT * K;
  z2* Y2 = Y + (size_t)BT * m * K;
  const size_t sV = (size_t)n * BT, sG = (size_t)BT * BT, sY = (size_t)BT * m;
  hipLaunchKernelGGL(k_bt_promote, dim3(2048, K), dim3(256), 0, s, (size_t)n * m, Z, ldZ, sZ, n,
                     m, C, ldC, sC);
  FORGE_LAUNCH();
  if (n < 2) return 0;
  rocblas_handle h = blas_handle(s);
  if (!h) return -1;
  const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0), mone(-1.0, 0.0);
  const i64 nblocks = (n - 1 + BT - 1) / BT;
  for (i64 bq = nblocks - 1; bq >= 0; --bq) {
    const i64 i0 = bq * BT;
    const int nb = (int)((n - 1 - i0) < BT ? (n - 1 - i0) : BT);
    const i64 mt = n - i0 - 1;
    hipLaunchKernelGGL(k_bt_vclean, dim3(cdiv(mt * nb, 256) < 4096 ? cdiv(mt * nb, 256) : 4096, K),
                       dim3(256), 0, s, n, i0, nb, t.A, t.sA, t.lda, V, sV);
    FORGE_LAUNCH();
    // G = V^H V
    RB(rocblas_zgemm_strided_batched(h, rocblas_operation_conjugate_transpose, rocblas_operation_none,
                                     nb, nb, (rocblas_int)mt, &one, (const rocblas_double_complex*)V,
                                     (rocblas_int)n, sV, (const rocblas_double_complex*)V,
                                     (rocblas_int)n, sV, &zero, (rocblas_double_complex*)G, BT, sG,
                                     K));
    hipLaunchKernelGGL(k_bt_tbuild, dim3(K), dim3(128), 0, s, nb, i0, n, G, t.tau, T);
    FORGE_LAUNCH();
    // Y = V^H C(i0+1:n, :)
    RB(rocblas_zgemm_strided_batched(h, rocblas_operation_conjugate_transpose, rocblas_operation_none,
                                     nb, (rocblas_int)m, (rocblas_int)mt, &one,
                                     (const rocblas_double_complex*)V, (rocblas_int)n, sV,
                                     (const rocblas_double_complex*)(C + i0 + 1), (rocblas_int)ldC,
                                     sC, &zero, (rocblas_double_complex*)Y, BT, sY, K));
    // Y2 = T Y
    RB(rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, nb,
                                     (rocblas_int)m, nb, &one, (const rocblas_double_complex*)T, BT,
                                     sG, (const rocblas_double_complex*)Y, BT, sY, &zero,
                                     (rocblas_double_complex*)Y2, BT, sY, K));
    // C(i0+1:n, :) -= V Y2
    RB(rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none,
                                     (rocblas_int)mt, (rocblas_int)m, nb, &mone,
                                     (const rocblas_double_complex*)V, (rocblas_int)n, sV,
                                     (const rocblas_double_complex*)Y2, BT, sY, &one,
                                     (rocblas_double_complex*)(C + i0 + 1), (rocblas_int)ldC, sC,
                                     K));
  }
  return 0;
}

}  // namespace forge
