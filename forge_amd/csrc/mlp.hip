// The MLP of the learned initial guess: 3 layers, width 128, ReLU,
//   Z1 = X W1 + b1, A1 = relu(Z1);  Z2 = A1 W2 + b2, A2 = relu(Z2);  Y = A2 W3 + b3
// with row-major activations (sample-major) and row-major weights (in x out).
// Parameters are packed [W1 (din x 128), b1, W2 (128 x 128), b2, W3 (128 x dout), b3].
//
// Training (FP32): forward, masked mean-squared loss, backward and Adam are
// hand-written kernels; the weight gradients are split-K products reduced in
// a fixed order, so a run is bitwise reproducible.  Inference: bf16 operands
// on the matrix cores (v_mfma_f32_16x16x16_bf16) with FP32 accumulation, one
// wave per 16 samples, activations passed between layers through LDS.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "kernels.hpp"

#define CK(x)               \
  do {                      \
    int _r = (x);           \
    if (_r != 0) return _r; \
  } while (0)

namespace forge {

constexpr int MLP_H = 128;
constexpr int MLP_DIN_MAX = 32;
constexpr int MLP_DOUT_MAX = 16;

struct MlpShape {
  i64 din, dout;
  i64 oW1, ob1, oW2, ob2, oW3, ob3, total;
};

static MlpShape mlp_shape(i64 din, i64 dout) {
  MlpShape s;
  s.din = din;
  s.dout = dout;
  s.oW1 = 0;
  s.ob1 = s.oW1 + din * MLP_H;
  s.oW2 = s.ob1 + MLP_H;
  s.ob2 = s.oW2 + MLP_H * MLP_H;
  s.oW3 = s.ob2 + MLP_H;
  s.ob3 = s.oW3 + MLP_H * dout;
  s.total = s.ob3 + dout;
  return s;
}

// ---------------------------------------------------------------- FP32 GEMM
// C (M x N, row-major, ldc = N) = op(A) op(B) over k in [k0, k1), written to
// C + z * M * N for split z.  op(A)(m, k) = A[m * K + k] (ta = 0) or A[k * M + m]
// (ta = 1); op(B)(k, n) = B[k * N + n] (tb = 0) or B[n * K + k] (tb = 1).
// 64 x 64 tiles, 256 threads with 4 x 4 outputs, k staged 16 at a time.
__global__ void __launch_bounds__(256) k_mlp_gemm(i64 M, i64 N, i64 K, const float* A, int ta,
                                                  const float* B, int tb, float* C, i64 kchunk) {
  __shared__ float As[16][64 + 1], Bs[16][64 + 1];
  const i64 m0 = (i64)blockIdx.x * 64, n0 = (i64)blockIdx.y * 64;
  const i64 k0 = (i64)blockIdx.z * kchunk;
  const i64 k1 = k0 + kchunk < K ? k0 + kchunk : K;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  float acc[4][4] = {};
  for (i64 kb = k0; kb < k1; kb += 16) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int kk = e >> 6, mm = e & 63;
      const i64 k = kb + kk, m = m0 + mm, nn = n0 + mm;
      float a = 0.f, b = 0.f;
      if (k < k1 && m < M) a = ta ? A[k * M + m] : A[m * K + k];
      if (k < k1 && nn < N) b = tb ? B[nn * K + k] : B[k * N + nn];
      As[kk][mm] = a;
      Bs[kk][mm] = b;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      float av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        av[u] = As[kk][ty + 16 * u];
        bv[u] = Bs[kk][tx + 16 * u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] += av[u] * bv[v];
    }
    __syncthreads();
  }
  float* Cz = C + (size_t)blockIdx.z * M * N;
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const i64 m = m0 + ty + 16 * u, nn = n0 + tx + 16 * v;
      if (m < M && nn < N) Cz[m * N + nn] = acc[u][v];
    }
}

// out[i] = sum_z part[z * len + i] (fixed order)
__global__ void k_mlp_sum_splits(i64 len, int S, const float* part, float* out) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= len) return;
  float s = 0.f;
  for (int z = 0; z < S; ++z) s += part[(size_t)z * len + i];
  out[i] = s;
}

// Z (B x N) += bias; A = relu(Z) (A may alias nothing; Z keeps the pre-activation)
__global__ void k_mlp_bias_act(i64 B, i64 N, float* Z, const float* bias, float* A, int relu) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= B * N) return;
  const float z = Z[i] + bias[i % N];
  Z[i] = z;
  if (A) A[i] = relu ? (z > 0.f ? z : 0.f) : z;
}

// dZ = dA * (Z > 0)
__global__ void k_mlp_relu_back(i64 len, const float* Z, const float* dA, float* dZ) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= len) return;
  dZ[i] = Z[i] > 0.f ? dA[i] : 0.f;
}

// per-block partials of sum mask (Y - T)^2 and of sum mask; dY = 2 mask (Y - T)
// (scaled by 1 / sum mask afterwards)
__global__ void k_mlp_loss(i64 len, const float* Y, const float* T, const float* mask, float* dY,
                           float* part) {
  __shared__ float s1[256], s2[256];
  float a = 0.f, b = 0.f;
  for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < len; i += (i64)gridDim.x * 256) {
    const float d = Y[i] - T[i], w = mask[i];
    a += w * d * d;
    b += w;
    dY[i] = 2.f * w * d;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      s1[threadIdx.x] += s1[threadIdx.x + s];
      s2[threadIdx.x] += s2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}

// loss = sum / count; dY *= 1 / count  (count from the partials, one block)
__global__ void k_mlp_loss_fin(i64 len, int nblk, const float* part, float* dY, float* loss) {
  __shared__ float inv;
  if (threadIdx.x == 0) {
    float a = 0.f, b = 0.f;
    for (int q = 0; q < nblk; ++q) {
      a += part[2 * q];
      b += part[2 * q + 1];
    }
    inv = b > 0.f ? 1.f / b : 0.f;
    *loss = a * inv;
  }
  __syncthreads();
  for (i64 i = threadIdx.x; i < len; i += 256) dY[i] *= inv;
}

// column sums of dZ (B x N) into db (N), one thread per column, fixed order
__global__ void k_mlp_colsum(i64 B, i64 N, const float* dZ, float* db) {
  const i64 c = (i64)blockIdx.x * 256 + threadIdx.x;
  if (c >= N) return;
  float s = 0.f;
  for (i64 r = 0; r < B; ++r) s += dZ[r * N + c];
  db[c] = s;
}

// Adam (bias-corrected), t >= 1
__global__ void k_mlp_adam(i64 len, float* p, const float* g, float* m, float* v, float lr,
                           float b1, float b2, float eps, float c1, float c2) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= len) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// rows idx[r] of src (rows of width w) into dst
__global__ void k_mlp_gather(i64 B, i64 w, const int* idx, const float* src, float* dst) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= B * w) return;
  dst[i] = src[(i64)idx[i / w] * w + i % w];
}

// ---------------------------------------------------------------- training state
struct MlpWork {
  MlpShape s;
  i64 Bmax = 0;
  int S = 0;  // split-K factor of the weight gradients
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  float *X = nullptr, *T = nullptr, *M = nullptr;  // batch
  float *Z1 = nullptr, *A1 = nullptr, *Z2 = nullptr, *A2 = nullptr, *Y = nullptr;
  float *dY = nullptr, *dA = nullptr, *dZ = nullptr, *dZ1 = nullptr, *part = nullptr;
  float* loss = nullptr;
};

static int mlp_alloc(MlpWork& w, i64 din, i64 dout, i64 B) {
  w.s = mlp_shape(din, dout);
  w.Bmax = B;
  w.S = (int)std::min<i64>(16, std::max<i64>(1, B / 512));
  const i64 P = w.s.total;
  float** bufs[] = {&w.p, &w.g, &w.m, &w.v};
  for (float** q : bufs) FORGE_HIP(hipMalloc((void**)q, sizeof(float) * P));
  FORGE_HIP(hipMalloc((void**)&w.X, sizeof(float) * B * din));
  FORGE_HIP(hipMalloc((void**)&w.T, sizeof(float) * B * dout));
  FORGE_HIP(hipMalloc((void**)&w.M, sizeof(float) * B * dout));
  float** act[] = {&w.Z1, &w.A1, &w.Z2, &w.A2, &w.dA, &w.dZ, &w.dZ1};
  for (float** q : act) FORGE_HIP(hipMalloc((void**)q, sizeof(float) * B * MLP_H));
  FORGE_HIP(hipMalloc((void**)&w.Y, sizeof(float) * B * dout));
  FORGE_HIP(hipMalloc((void**)&w.dY, sizeof(float) * B * dout));
  FORGE_HIP(hipMalloc((void**)&w.part, sizeof(float) * std::max<i64>(w.S * MLP_H * MLP_H, 2 * 1024)));
  FORGE_HIP(hipMalloc((void**)&w.loss, sizeof(float) * 4));
  return 0;
}

static void mlp_free(MlpWork& w) {
  float* all[] = {w.p, w.g, w.m, w.v, w.X, w.T, w.M, w.Z1, w.A1, w.Z2, w.A2, w.Y, w.dY,
                  w.dA, w.dZ, w.dZ1, w.part, w.loss};
  for (float* q : all)
    if (q) (void)hipFree(q);
  w = MlpWork{};
}

static dim3 gemm_grid(i64 M, i64 N, int S) {
  return dim3((unsigned)cdiv(M, 64), (unsigned)cdiv(N, 64), (unsigned)S);
}

// C = op(A) op(B), the K range split S ways when S > 1 (partials summed in order)
static int gemm(i64 M, i64 N, i64 K, const float* A, int ta, const float* B, int tb, float* C,
                int S, float* part, hipStream_t st) {
  if (S <= 1) {
    hipLaunchKernelGGL(k_mlp_gemm, gemm_grid(M, N, 1), dim3(256), 0, st, M, N, K, A, ta, B, tb, C, K);
    FORGE_LAUNCH();
    return 0;
  }
  const i64 kc = (cdiv(K, S) + 15) / 16 * 16;
  const int Sx = (int)cdiv(K, kc);
  hipLaunchKernelGGL(k_mlp_gemm, gemm_grid(M, N, Sx), dim3(256), 0, st, M, N, K, A, ta, B, tb, part, kc);
  FORGE_LAUNCH();
  hipLaunchKernelGGL(k_mlp_sum_splits, dim3((unsigned)cdiv(M * N, 256)), dim3(256), 0, st, M * N, Sx,
                     part, C);
  FORGE_LAUNCH();
  return 0;
}

// forward of B rows already in w.X; Y = network output
static int forward(MlpWork& w, i64 B, hipStream_t st) {
  const MlpShape& s = w.s;
  const float* p = w.p;
  CK(gemm(B, MLP_H, s.din, w.X, 0, p + s.oW1, 0, w.Z1, 1, nullptr, st));
  hipLaunchKernelGGL(k_mlp_bias_act, dim3((unsigned)cdiv(B * MLP_H, 256)), dim3(256), 0, st, B,
                     (i64)MLP_H, w.Z1, p + s.ob1, w.A1, 1);
  FORGE_LAUNCH();
  CK(gemm(B, MLP_H, MLP_H, w.A1, 0, p + s.oW2, 0, w.Z2, 1, nullptr, st));
  hipLaunchKernelGGL(k_mlp_bias_act, dim3((unsigned)cdiv(B * MLP_H, 256)), dim3(256), 0, st, B,
                     (i64)MLP_H, w.Z2, p + s.ob2, w.A2, 1);
  FORGE_LAUNCH();
  CK(gemm(B, s.dout, MLP_H, w.A2, 0, p + s.oW3, 0, w.Y, 1, nullptr, st));
  hipLaunchKernelGGL(k_mlp_bias_act, dim3((unsigned)cdiv(B * s.dout, 256)), dim3(256), 0, st, B,
                     s.dout, w.Y, p + s.ob3, (float*)nullptr, 0);
  FORGE_LAUNCH();
  return 0;
}

// loss and gradient (into w.g) of the B rows in w.X / w.T / w.M
static int forward_backward(MlpWork& w, i64 B, hipStream_t st) {
  const MlpShape& s = w.s;
  const float* p = w.p;
  float* g = w.g;
  CK(forward(w, B, st));
  const i64 ly = B * s.dout;
  const int nblk = (int)std::min<i64>(1024, cdiv(ly, 256));
  hipLaunchKernelGGL(k_mlp_loss, dim3(nblk), dim3(256), 0, st, ly, w.Y, w.T, w.M, w.dY, w.part);
  FORGE_LAUNCH();
  hipLaunchKernelGGL(k_mlp_loss_fin, dim3(1), dim3(256), 0, st, ly, nblk, w.part, w.dY, w.loss);
  FORGE_LAUNCH();
  // layer 3
  CK(gemm(MLP_H, s.dout, B, w.A2, 1, w.dY, 0, g + s.oW3, w.S, w.part, st));
  hipLaunchKernelGGL(k_mlp_colsum, dim3((unsigned)cdiv(s.dout, 256)), dim3(256), 0, st, B, s.dout,
                     w.dY, g + s.ob3);
  FORGE_LAUNCH();
  CK(gemm(B, MLP_H, s.dout, w.dY, 0, p + s.oW3, 1, w.dA, 1, nullptr, st));
  hipLaunchKernelGGL(k_mlp_relu_back, dim3((unsigned)cdiv(B * MLP_H, 256)), dim3(256), 0, st,
                     B * MLP_H, w.Z2, w.dA, w.dZ);
  FORGE_LAUNCH();
  // layer 2
  CK(gemm(MLP_H, MLP_H, B, w.A1, 1, w.dZ, 0, g + s.oW2, w.S, w.part, st));
  hipLaunchKernelGGL(k_mlp_colsum, dim3(1), dim3(256), 0, st, B, (i64)MLP_H, w.dZ, g + s.ob2);
  FORGE_LAUNCH();
  CK(gemm(B, MLP_H, MLP_H, w.dZ, 0, p + s.oW2, 1, w.dA, 1, nullptr, st));
  hipLaunchKernelGGL(k_mlp_relu_back, dim3((unsigned)cdiv(B * MLP_H, 256)), dim3(256), 0, st,
                     B * MLP_H, w.Z1, w.dA, w.dZ1);
  FORGE_LAUNCH();
  // layer 1
  CK(gemm(s.din, MLP_H, B, w.X, 1, w.dZ1, 0, g + s.oW1, w.S, w.part, st));
  hipLaunchKernelGGL(k_mlp_colsum, dim3(1), dim3(256), 0, st, B, (i64)MLP_H, w.dZ1, g + s.ob1);
  FORGE_LAUNCH();
  return 0;
}

// ---------------------------------------------------------------- bf16 inference
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline short to_bf16(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// Weights of one layer in bf16, laid out as the B operand fragments:
// Wb[(kt * ntiles + nt) * 64 + lane] holds W[16 kt + 4 (lane / 16) + i][16 nt + lane % 16], i = 0..3.
__global__ void k_mlp_pack_b(i64 K, i64 N, i64 Kp, i64 Np, const float* W, bf16x4* Wb) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  const i64 ntl = Np / 16, total = (Kp / 16) * ntl * 64;
  if (e >= total) return;
  const int lane = (int)(e & 63);
  const i64 t = e >> 6, kt = t / ntl, nt = t % ntl;
  bf16x4 r;
  for (int i = 0; i < 4; ++i) {
    const i64 k = 16 * kt + 4 * (lane / 16) + i, n = 16 * nt + lane % 16;
    r[i] = to_bf16((k < K && n < N) ? W[k * N + n] : 0.f);
  }
  Wb[e] = r;
}

// One wave per 16 samples (4 waves per block).  Per layer: the A operand
// (16 samples x 16 k) is read from the wave's LDS rows, D (16 x 16 per
// column tile) gets bias + ReLU in FP32 and goes back to LDS as bf16.
// MFMA 16x16x16 lane layout: A: row lane % 16, k 4 (lane / 16) + i;
// B: k 4 (lane / 16) + i, column lane % 16; D: row 4 (lane / 16) + i, column lane % 16.
__global__ void __launch_bounds__(256) k_mlp_infer_bf16(i64 N, i64 din, i64 dout, const float* X,
                                                        const bf16x4* W1b, const bf16x4* W2b,
                                                        const bf16x4* W3b, const float* b1,
                                                        const float* b2, const float* b3,
                                                        float* Y) {
  __shared__ short act[4][16][MLP_H + 8];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const i64 r0 = ((i64)blockIdx.x * 4 + wv) * 16;
  if (r0 >= N) return;  // whole wave leaves together; no block barrier below
  const int rl = lane % 16, kq = lane / 16;
  // input rows into LDS (bf16), din padded to 32
  for (int e = lane; e < 16 * 32; e += 64) {
    const int rr = e / 32, c = e % 32;
    const i64 r = r0 + rr;
    act[wv][rr][c] = to_bf16((r < N && c < din) ? X[r * din + c] : 0.f);
  }
  __builtin_amdgcn_wave_barrier();
  auto layer = [&](int ktiles, int ntiles, const bf16x4* Wb, const float* bias, int nvalid,
                   bool relu, float* out) {
    f32x4 acc[MLP_H / 16];
    for (int nt = 0; nt < ntiles; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int kt = 0; kt < ktiles; ++kt) {
      bf16x4 a;
      for (int i = 0; i < 4; ++i) a[i] = act[wv][rl][16 * kt + 4 * kq + i];
      for (int nt = 0; nt < ntiles; ++nt)
        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, Wb[(kt * ntiles + nt) * 64 + lane],
                                                             acc[nt], 0, 0, 0);
    }
    __builtin_amdgcn_wave_barrier();
    for (int nt = 0; nt < ntiles; ++nt) {
      const int c = 16 * nt + rl;
      for (int i = 0; i < 4; ++i) {
        const int rr = 4 * kq + i;
        float z = acc[nt][i] + (c < nvalid ? bias[c] : 0.f);
        if (relu) z = z > 0.f ? z : 0.f;
        if (out) {
          const i64 r = r0 + rr;
          if (r < N && c < nvalid) out[r * nvalid + c] = z;
        } else {
          act[wv][rr][c] = to_bf16(z);
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  };
  layer(2, MLP_H / 16, W1b, b1, MLP_H, true, nullptr);
  layer(MLP_H / 16, MLP_H / 16, W2b, b2, MLP_H, true, nullptr);
  layer(MLP_H / 16, 1, W3b, b3, (int)dout, false, Y);
}

}  // namespace forge

using namespace forge;

static int check_dims(i64 din, i64 dout, const char* who) {
  if (din < 1 || din > MLP_DIN_MAX || dout < 1 || dout > MLP_DOUT_MAX) {
    set_error(std::string(who) + ": din must be in 1..32 and dout in 1..16");
    return -2;
  }
  return 0;
}

extern "C" int64_t forge_mlp_param_count(int64_t din, int64_t dout) {
  return mlp_shape(din, dout).total;
}

extern "C" int forge_mlp_loss_grad(const float* params, int64_t din, int64_t dout, const float* X,
                                   const float* T, const float* mask, int64_t N, float* loss,
                                   float* grad) {
  if (!params || !X || !T || !mask || !loss || !grad || N < 1) {
    set_error("forge_mlp_loss_grad: invalid argument");
    return -2;
  }
  CK(check_dims(din, dout, "forge_mlp_loss_grad"));
  MlpWork w;
  int rc = mlp_alloc(w, din, dout, N);
  hipStream_t st = nullptr;
  if (rc == 0) {
    const i64 P = w.s.total;
    rc = [&]() -> int {
      FORGE_HIP(hipMemcpy(w.p, params, sizeof(float) * P, hipMemcpyHostToDevice));
      FORGE_HIP(hipMemcpy(w.X, X, sizeof(float) * N * din, hipMemcpyHostToDevice));
      FORGE_HIP(hipMemcpy(w.T, T, sizeof(float) * N * dout, hipMemcpyHostToDevice));
      FORGE_HIP(hipMemcpy(w.M, mask, sizeof(float) * N * dout, hipMemcpyHostToDevice));
      CK(forward_backward(w, N, st));
      FORGE_HIP(hipMemcpy(grad, w.g, sizeof(float) * P, hipMemcpyDeviceToHost));
      FORGE_HIP(hipMemcpy(loss, w.loss, sizeof(float), hipMemcpyDeviceToHost));
      return 0;
    }();
  }
  mlp_free(w);
  return rc;
}

// splitmix64, for the reproducible batch order
static uint64_t smix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

extern "C" int forge_mlp_train(float* params, int64_t din, int64_t dout, const float* X,
                               const float* T, const float* mask, int64_t N, int64_t steps,
                               int64_t batch, double lr, uint64_t seed, float* loss_hist) {
  if (!params || !X || !T || !mask || N < 1 || steps < 0 || batch < 1) {
    set_error("forge_mlp_train: invalid argument");
    return -2;
  }
  CK(check_dims(din, dout, "forge_mlp_train"));
  const i64 B = std::min<i64>(batch, N);
  MlpWork w;
  float *Xd = nullptr, *Td = nullptr, *Md = nullptr;
  int* idx = nullptr;
  hipStream_t st = nullptr;
  int rc = mlp_alloc(w, din, dout, B);
  if (rc == 0) {
    rc = [&]() -> int {
      const i64 P = w.s.total;
      FORGE_HIP(hipMalloc((void**)&Xd, sizeof(float) * N * din));
      FORGE_HIP(hipMalloc((void**)&Td, sizeof(float) * N * dout));
      FORGE_HIP(hipMalloc((void**)&Md, sizeof(float) * N * dout));
      FORGE_HIP(hipMalloc((void**)&idx, sizeof(int) * B));
      FORGE_HIP(hipMemcpy(Xd, X, sizeof(float) * N * din, hipMemcpyHostToDevice));
      FORGE_HIP(hipMemcpy(Td, T, sizeof(float) * N * dout, hipMemcpyHostToDevice));
      FORGE_HIP(hipMemcpy(Md, mask, sizeof(float) * N * dout, hipMemcpyHostToDevice));
      FORGE_HIP(hipMemcpy(w.p, params, sizeof(float) * P, hipMemcpyHostToDevice));
      FORGE_HIP(hipMemset(w.m, 0, sizeof(float) * P));
      FORGE_HIP(hipMemset(w.v, 0, sizeof(float) * P));
      std::vector<int> perm(N);
      for (i64 i = 0; i < N; ++i) perm[i] = (int)i;
      uint64_t s = seed;
      i64 pos = N;  // forces a shuffle on the first step
      const float b1 = 0.9f, b2 = 0.999f;
      float c1 = 1.f, c2 = 1.f;
      for (i64 t = 1; t <= steps; ++t) {
        if (pos + B > N) {  // new epoch: Fisher-Yates shuffle
          for (i64 i = N - 1; i > 0; --i) std::swap(perm[i], perm[smix(s) % (uint64_t)(i + 1)]);
          pos = 0;
        }
        FORGE_HIP(hipMemcpyAsync(idx, perm.data() + pos, sizeof(int) * B, hipMemcpyHostToDevice, st));
        pos += B;
        hipLaunchKernelGGL(k_mlp_gather, dim3((unsigned)cdiv(B * din, 256)), dim3(256), 0, st, B, din,
                           idx, Xd, w.X);
        hipLaunchKernelGGL(k_mlp_gather, dim3((unsigned)cdiv(B * dout, 256)), dim3(256), 0, st, B,
                           dout, idx, Td, w.T);
        hipLaunchKernelGGL(k_mlp_gather, dim3((unsigned)cdiv(B * dout, 256)), dim3(256), 0, st, B,
                           dout, idx, Md, w.M);
        FORGE_LAUNCH();
        CK(forward_backward(w, B, st));
        c1 *= b1;
        c2 *= b2;
        hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)cdiv(P, 256)), dim3(256), 0, st, P, w.p, w.g,
                           w.m, w.v, (float)lr, b1, b2, 1e-8f, 1.f - c1, 1.f - c2);
        FORGE_LAUNCH();
        if (loss_hist) {
          FORGE_HIP(hipMemcpyAsync(loss_hist + t - 1, w.loss, sizeof(float), hipMemcpyDeviceToHost, st));
        }
        // the host vector perm is rewritten by the next shuffle: wait for the copy
        if (pos + B > N) FORGE_HIP(hipStreamSynchronize(st));
      }
      FORGE_HIP(hipStreamSynchronize(st));
      FORGE_HIP(hipMemcpy(params, w.p, sizeof(float) * P, hipMemcpyDeviceToHost));
      return 0;
    }();
  }
  if (Xd) (void)hipFree(Xd);
  if (Td) (void)hipFree(Td);
  if (Md) (void)hipFree(Md);
  if (idx) (void)hipFree(idx);
  mlp_free(w);
  return rc;
}

// FP32 forward (the training kernels) or bf16 matrix-core inference
extern "C" int forge_mlp_infer(const float* params, int64_t din, int64_t dout, const float* X,
                               int64_t N, int32_t bf16, float* Y) {
  if (!params || !X || !Y || N < 1) {
    set_error("forge_mlp_infer: invalid argument");
    return -2;
  }
  CK(check_dims(din, dout, "forge_mlp_infer"));
  const MlpShape s = mlp_shape(din, dout);
  if (!bf16) {
    MlpWork w;
    int rc = mlp_alloc(w, din, dout, N);
    if (rc == 0) {
      rc = [&]() -> int {
        FORGE_HIP(hipMemcpy(w.p, params, sizeof(float) * s.total, hipMemcpyHostToDevice));
        FORGE_HIP(hipMemcpy(w.X, X, sizeof(float) * N * din, hipMemcpyHostToDevice));
        CK(forward(w, N, nullptr));
        FORGE_HIP(hipMemcpy(Y, w.Y, sizeof(float) * N * dout, hipMemcpyDeviceToHost));
        return 0;
      }();
    }
    mlp_free(w);
    return rc;
  }
  float *p = nullptr, *Xd = nullptr, *Yd = nullptr;
  bf16x4 *W1b = nullptr, *W2b = nullptr, *W3b = nullptr;
  const i64 n1 = (32 / 16) * (MLP_H / 16) * 64, n2 = (MLP_H / 16) * (MLP_H / 16) * 64,
            n3 = (MLP_H / 16) * 1 * 64;
  int rc = [&]() -> int {
    FORGE_HIP(hipMalloc((void**)&p, sizeof(float) * s.total));
    FORGE_HIP(hipMalloc((void**)&Xd, sizeof(float) * N * din));
    FORGE_HIP(hipMalloc((void**)&Yd, sizeof(float) * N * dout));
    FORGE_HIP(hipMalloc((void**)&W1b, sizeof(bf16x4) * n1));
    FORGE_HIP(hipMalloc((void**)&W2b, sizeof(bf16x4) * n2));
    FORGE_HIP(hipMalloc((void**)&W3b, sizeof(bf16x4) * n3));
    FORGE_HIP(hipMemcpy(p, params, sizeof(float) * s.total, hipMemcpyHostToDevice));
    FORGE_HIP(hipMemcpy(Xd, X, sizeof(float) * N * din, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_pack_b, dim3((unsigned)cdiv(n1, 256)), dim3(256), 0, 0, din, (i64)MLP_H,
                       (i64)32, (i64)MLP_H, p + s.oW1, W1b);
    hipLaunchKernelGGL(k_mlp_pack_b, dim3((unsigned)cdiv(n2, 256)), dim3(256), 0, 0, (i64)MLP_H,
                       (i64)MLP_H, (i64)MLP_H, (i64)MLP_H, p + s.oW2, W2b);
    hipLaunchKernelGGL(k_mlp_pack_b, dim3((unsigned)cdiv(n3, 256)), dim3(256), 0, 0, (i64)MLP_H, dout,
                       (i64)MLP_H, (i64)16, p + s.oW3, W3b);
    FORGE_LAUNCH();
    hipLaunchKernelGGL(k_mlp_infer_bf16, dim3((unsigned)cdiv(N, 64)), dim3(256), 0, 0, (i64)N, din,
                       dout, Xd, W1b, W2b, W3b, p + s.ob1, p + s.ob2, p + s.ob3, Yd);
    FORGE_LAUNCH();
    FORGE_HIP(hipMemcpy(Y, Yd, sizeof(float) * N * dout, hipMemcpyDeviceToHost));
    return 0;
  }();
  void* all[] = {p, Xd, Yd, W1b, W2b, W3b};
  for (void* q : all)
    if (q) (void)hipFree(q);
  return rc;
}
