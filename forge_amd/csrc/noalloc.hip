// zdiagRed for callers that cannot allocate device memory.
//
// forge_zdiagred_noalloc is forge_zdiagred (LapackRoutines.f90:381-433) with
// every device buffer of the eigensolver -- the matrix, the Householder
// workspace, the divide-and-conquer workspace, the tridiagonal eigenvectors,
// the eigenvalues, the complex eigenvectors and the back-transformation
// workspace -- carved out of one file-scope device array, `scratch`, of
// SCRATCH_BYTES per device.  Orders up to NOALLOC_NMAX fit.  The call makes
// no hipMalloc, hipFree, stream or event creation and no library call that
// allocates; it runs on the null stream, and the back-transformation uses a
// fixed block width instead of the tuner (which times with events of its
// own).
//
// Each device holds its own copy of `scratch` for the life of the process.
// Its address is looked up once per device (hipGetSymbolAddress, with that
// device current) on the first call there and kept; host code uses only that
// address, never the array's name.  Every call on a device shares that
// device's array: calls from several host threads are serialised by a lock,
// and calls that overlap on one device through other means (another
// process's copy of the library does not share it) race.  The array keeps
// the last call's contents; no kernel relies on it starting at zero: the
// divide-and-conquer clears its own workspace on every call, and every other
// buffer is written before it is read.
#include <mutex>

#include "kernels.hpp"

namespace forge {

constexpr size_t SCRATCH_BYTES = (size_t)32 << 20;  // 32 MiB per device
constexpr i64 NOALLOC_NMAX = 512;
constexpr int NOALLOC_MAXDEV = 64;

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

static std::mutex g_noalloc_mu;
static void* g_scratch_addr[NOALLOC_MAXDEV] = {};

// device address of this device's scratch (looked up once per device)
static int scratch_address(void** out) {
  int dev = 0;
  FORGE_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAXDEV) {
    set_error("forge_zdiagred_noalloc: device index out of range");
    return -3;
  }
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("forge_zdiagred_noalloc: hipGetSymbolAddress(scratch): ") +
                hipGetErrorString(e));
      return -1;
    }
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return 0;
}

struct NoallocLayout {
  size_t H, trd, dc, Zq, ev, Cv, bt, total;
};

static NoallocLayout noalloc_layout(i64 n, i64 iu) {
  NoallocLayout L;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const i64 ld = (n + 15) & ~(i64)15;
  L.H = take(sizeof(z2) * ld * n);
  L.trd = take(tridiag_workspace_bytes(n, 1));
  L.dc = take(dc_workspace_bytes(n, 1, iu));
  L.Zq = take(sizeof(double) * n * iu);
  L.ev = take(sizeof(double) * n);
  L.Cv = take(sizeof(z2) * n * iu);
  L.bt = take(backtr_workspace_bytes(n, 1, iu));
  L.total = off;
  return L;
}

}  // namespace forge

using namespace forge;

extern "C" int forge_zdiagred_noalloc(forge_z* A, int64_t n, double* eigenvalues, int64_t il,
                                      int64_t iu, int32_t* info) {
  if (!A || !eigenvalues || n < 1 || il < 1 || iu < il || iu > n) {
    set_error("forge_zdiagred_noalloc: invalid argument");
    return -2;
  }
  if (n > NOALLOC_NMAX) {
    set_error("forge_zdiagred_noalloc: n = " + std::to_string(n) + " exceeds the maximum order " +
              std::to_string(NOALLOC_NMAX) + " of the preallocated device scratch (use forge_zdiagred)");
    return -2;
  }
  const NoallocLayout L = noalloc_layout(n, iu);
  if (L.total > SCRATCH_BYTES) {
    set_error("forge_zdiagred_noalloc: the workspace exceeds the preallocated device scratch");
    return -2;
  }
  if (forge_device_count() < 1) { set_error("forge_zdiagred_noalloc: no HIP device"); return -3; }
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  void* base = nullptr;
  {
    const int r = scratch_address(&base);
    if (r) return r;
  }
  char* b = (char*)base;
  const hipStream_t s = nullptr;  // the null stream: creating one could allocate
  const i64 ld = (n + 15) & ~(i64)15;
  z2* H = (z2*)(b + L.H);
  double* Zq = (double*)(b + L.Zq);
  double* ev = (double*)(b + L.ev);
  z2* Cv = (z2*)(b + L.Cv);
  TridiagWork t{};
  {
    const int r = tridiag_bind(t, b + L.trd, n, 1);
    if (r) return r;
  }
  t.A = H;
  t.sA = (size_t)ld * n;
  t.lda = ld;
  FORGE_HIP(hipMemcpy2DAsync(H, sizeof(z2) * ld, A, sizeof(z2) * n, sizeof(z2) * n, n,
                             hipMemcpyHostToDevice, s));
  {
    const int r = tridiag(t, s, nullptr, nullptr, nullptr);
    if (r) return r;
  }
  const int fails = dc_solve(t.d, t.e, n, 1, iu, ev, Zq, b + L.dc, s);
  if (fails < 0) return fails;
  {
    const int r = backtransform(t, Zq, n, (size_t)n * iu, iu, Cv, n, (size_t)n * iu, b + L.bt, s,
                                nullptr, nullptr, 128);
    if (r) return r;
  }
  const i64 m = iu - il + 1;
  FORGE_HIP(hipMemcpyAsync(A, Cv + n * (il - 1), sizeof(z2) * n * m, hipMemcpyDeviceToHost, s));
  FORGE_HIP(hipMemcpyAsync(eigenvalues, ev + (il - 1), sizeof(double) * m, hipMemcpyDeviceToHost, s));
  FORGE_HIP(hipStreamSynchronize(s));
  if (info) *info = fails;
  return fails;
}
