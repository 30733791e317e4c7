// C ABI of libforge_hip: context, reference-interface entry points and the
// device-resident SCF loop (the call sequence of Main.f90:236-679, numS = 1).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace forge {

static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }

static const double PI = 3.14159265358979323846264338327950288;

// small kernels local to the engine -----------------------------------------
__global__ void k_zero_upper(i64 n, i64 ld, z2* H) {
  i64 q = (i64)blockIdx.x * 256 + threadIdx.x;
  if (q >= n * n) return;
  i64 r = q % n, c = q / n;
  if (r < c) H[r + ld * c] = mk(0.0, 0.0);
}

// leading dimension of the eigensolver's matrices: a multiple of 16 elements
// keeps every 16-row column chunk one aligned 256-byte segment
static inline i64 pad_ld(i64 n) { return (n + 15) & ~(i64)15; }

// With ya = L a and yb = L b: out[0] = a.ya, out[1] = a.yb, out[2] = b.yb,
// out[3] = a.a, out[4] = a.b, out[5] = b.b (single block, fixed order, so
// the sums are reproducible); b and yb may be null (terms read 0).
__global__ void k_dots(i64 n, const double* a, const double* b, const double* ya,
                       const double* yb, double* out) {
  __shared__ double sh[6][256];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (i64 i = threadIdx.x; i < n; i += 256) {
    double av = a[i], bv = b ? b[i] : 0.0, yav = ya[i], ybv = yb ? yb[i] : 0.0;
    acc[0] += av * yav;
    acc[1] += av * ybv;
    acc[2] += bv * ybv;
    acc[3] += av * av;
    acc[4] += av * bv;
    acc[5] += bv * bv;
  }
  for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 6; ++q) out[q] = sh[q][0];
}

// bias term of GetKineticEnergy: sum_i (d_i + 1/2) bias_i for d = a and b.
__global__ void k_bias_dot(i64 n, const double* a, const double* b, const double* bias,
                           double* out) {
  __shared__ double sh[2][256];
  double s0 = 0, s1 = 0;
  for (i64 i = threadIdx.x; i < n; i += 256) {
    s0 += (a[i] + 0.5) * bias[i];
    if (b) s1 += (b[i] + 0.5) * bias[i];
  }
  sh[0][threadIdx.x] = s0;
  sh[1][threadIdx.x] = s1;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[0] = sh[0][0]; out[1] = sh[1][0]; }
}

// pot = 2 alphaH ((1-s) ya + s yb) + U ((1-s) da + s db); d_out = (1-s) da + s db
__global__ void k_potential(i64 n, const double* ya, const double* yb, const double* da,
                            const double* db, double s, double alphaH, double U, double* pot,
                            double* dout) {
  i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double y = yb ? (1.0 - s) * ya[i] + s * yb[i] : ya[i];
  double d = db ? (1.0 - s) * da[i] + s * db[i] : da[i];
  pot[i] = 2.0 * alphaH * y + U * d;
  if (dout) dout[i] = d;
}

// Psi(:, s) = Cwin(:, idx[s])
__global__ void k_gather_cols(i64 n, i64 S, const int* idx, const z2* Cwin, z2* Psi) {
  i64 q = (i64)blockIdx.x * 256 + threadIdx.x;
  if (q >= n * S) return;
  i64 r = q % n, s = q / n;
  Psi[q] = Cwin[r + n * (i64)idx[s]];
}

// window columns nL-1 .. nU-1 of C (batch) -> Cwin(:, b + nb*k)
__global__ void k_copy_window(i64 n, i64 nb, i64 col0, int K, i64 k0, const z2* C, size_t ldC,
                              size_t sC, z2* Cwin) {
  i64 q = (i64)blockIdx.x * 256 + threadIdx.x;
  int kb = blockIdx.y;
  if (q >= n * nb) return;
  i64 r = q % n, b = q / n;
  Cwin[r + n * (b + nb * (k0 + kb))] = C[kb * sC + r + ldC * (col0 + b)];
}

}  // namespace forge

using namespace forge;

// pinned readback entries beyond the eigenvalues and site densities (5 energy
// sums, 10 dot products)
constexpr i64 HPIN_SMALL = 64;

struct forge_ctx {
  forge_params P;
  PairParams pp;
  hipStream_t s = nullptr;
  i64 n = 0, nnc = 0, numk = 0, Nk = 0, nL = 0, nU = 0, nb = 0;
  std::vector<double> Coords, tn, MomVal;
  std::vector<i64> nMomComp, u1, u2;
  DeviceGeom g;
  double *X = nullptr, *V = nullptr, *T = nullptr, *L = nullptr, *bias = nullptr, *pot = nullptr;
  z2 *rho = nullptr, *rho_new = nullptr;  // rho_new has nnc+1 matrices (last = R scratch)
  // eigensolver batch
  int Kb = 0;
  i64 ldH = 0;
  z2* H = nullptr;
  void* trd_ws = nullptr;
  TridiagWork trd{};
  void* ts_ws = nullptr;  // two-stage reduction workspace (Kb matrices), when allocated
  TwoStageWork ts;
  void* dc_ws = nullptr;
  double *Zq = nullptr, *evals = nullptr;
  z2* Cv = nullptr;
  void* bt_ws = nullptr;
  z2* Cwin = nullptr;  // n x nb x Nk
  double *ph_asm = nullptr, *ph_cell = nullptr;  // device [Nk][nnc][2]
  std::vector<double> h_ph_asm, h_ph_cell;
  z2* Psi = nullptr;
  double* wph = nullptr;
  int* gidx = nullptr;
  double *part = nullptr, *e5 = nullptr, *dots = nullptr;
  double *da = nullptr, *db = nullptr, *ya = nullptr, *yb = nullptr;
  // FP32 copy of the eigensolver batch for the mixed-precision hemv
  // (allocated on the first forge_scf_set_options with precision 1)
  float2* S = nullptr;
  // pinned host memory for the readbacks of every iteration (eigenvalues
  // n x Kb, the two site densities, energy sums): the copies land by DMA
  // and are complete when the stream is synchronised
  double* hpin = nullptr;
  // SCF state
  forge_scf_options opt{};
  bool scf_ready = false;
  bool mixed_done = false;  // precision 1: switched to FP64 for the rest of the solve
  double e[4] = {0, 0, 0, 0}, Etot = 0;
  std::vector<double> h_da;
  int it = 0;
  double tm[FORGE_NTIMINGS] = {};
  hipEvent_t ev[16] = {};
  // C3-reduced k-set (nenforceC3 = 1): tables built by c3_configure from
  // RotateLayers and nC3pairs; c3 = the SCF loop uses Solve_C3 / GetFock_C3
  bool c3 = false, c3_ready = false;
  std::vector<int64_t> c3_key;
  std::vector<i64> c3_reps, c3_k2, c3_k3;  // representatives; per flat k its partners
  std::vector<char> c3_inv;                // per flat k: C3-invariant
  std::vector<C3Col> c3_expand, c3_psi;
  std::vector<C3Slot> c3_slots;            // per representative
  std::vector<double> c3_wph_h;
  C3Geom c3g{};
  int* c3_perm = nullptr;
  z2* c3_tab = nullptr;
  double* c3_ph_asm = nullptr;
  C3Col *c3_dexpand = nullptr, *c3_dpsi = nullptr;
  z2* c3_Psi = nullptr;
  double* c3_wph = nullptr;
  i64 c3_cap = 0;  // columns of c3_Psi
  // nenforceT / C2 / C2T of the SCF loop (bits 1, 2, 4) and the C2 partners
  int sym = 0;
  int* c2_perm = nullptr;
  // nenforceValley: the loop's window density is the mean of the plain and
  // the valley-projected states' (Main.f90:469-494)
  bool valley = false, valley_ready = false;
  ValleyGeom vg{};
  void* valley_ws = nullptr;
  std::vector<double> h_vk;  // [Nk][2]
  // the device-resident iteration (options.graph 2: eager, 3: one HIP graph
  // per phase, captured on first use and replayed)
  double* dband = nullptr;  // nb * Nk window energies, band-major
  i64* dS = nullptr;        // weighted window states
  double* docc = nullptr;   // E_F, nOcc, nPart, DegFactor
  double* dst = nullptr;    // [0..3] energies of the state, [4] E_tot, [5] step
  double* drow = nullptr;   // trajectory row of the iteration
  bool dev_iter = false;    // eig_batch: reduction without graph or syncs, deferred D&C
  bool dev_state = false;   // dst holds the state's energies
  bool dev_warm[2] = {false, false};  // a host-path iteration of the phase ran
  hipGraphExec_t it_graph[2] = {nullptr, nullptr};  // mixed phase, FP64
};

static double total_energy(const forge_params& P, const double* e) {
  return scf_total_energy(P.alpha, P.alphaH, P.U, e);
}

extern "C" const char* forge_last_error(void) { return g_err.c_str(); }

extern "C" int forge_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

#define CK(x)                 \
  do {                        \
    int _r = (x);             \
    if (_r != 0) return _r;   \
  } while (0)

static int dalloc(void** p, size_t bytes) {
  FORGE_HIP(hipMalloc(p, bytes ? bytes : 256));
  return 0;
}

// Reduction to tridiagonal form: 1 one-stage (tridiag.hip), 2 two-stage
// (twostage.hip).  FORGE_REDUCTION=1 or 2 in the environment overrides the
// request; a request of 0 picks two-stage from order TS_MIN_N on, where its
// matrix-core products outrun the one-stage reduction's matrix-vector
// products.
constexpr i64 TS_MIN_N = 1024;
static int reduction_mode(int requested, i64 n) {
  const char* env = getenv("FORGE_REDUCTION");
  if (env && (env[0] == '1' || env[0] == '2') && env[1] == 0) return env[0] - '0';
  if (requested == 1 || requested == 2) return requested;
  return n >= TS_MIN_N ? 2 : 1;
}

extern "C" int forge_ctx_create(const forge_params* p, const double* Coords, const double* tn,
                                const double* MomentaValues, const int64_t* nMomentaComponents,
                                const int64_t* nUnitCell_1, const int64_t* nUnitCell_2,
                                forge_ctx** out) {
  if (!p || !Coords || !tn || !MomentaValues || !nMomentaComponents || !nUnitCell_1 ||
      !nUnitCell_2 || !out) {
    set_error("forge_ctx_create: null argument");
    return -2;
  }
  if (p->ndim < 2 || p->nnc < 1 || p->nnc > MAXCELL || p->numk < 1 || p->nUpper > p->ndim ||
      p->nLower < 1 || p->nLower > p->nUpper) {
    set_error("forge_ctx_create: invalid sizes in forge_params");
    return -2;
  }
  if (p->fphase != 0) {
    set_error("forge_ctx_create: fphase != 0 is not supported (Setup.f90:25 uses 0)");
    return -2;
  }
  {
    // the Fermi block reads sorted window states nF and nF + 1 (Main.f90:400-401)
    const i64 Nk = p->numk * p->numk, nbw = p->nUpper - p->nLower + 1;
    const double x = (double)(Nk * ((p->ndim / 2 - p->nLower) + 1)) + (double)Nk / 2.0 * p->nfilling;
    const i64 nF = (i64)(x >= 0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5));
    if (nF < 1 || nF > nbw * Nk - 1) {
      set_error("forge_ctx_create: the filling puts the Fermi level outside the band window "
                "(need 1 <= nF <= (nUpper - nLower + 1) numk^2 - 1)");
      return -2;
    }
  }
  int ndev = forge_device_count();
  if (ndev < 1) {
    set_error("forge_ctx_create: no HIP device");
    return -3;
  }
  forge_ctx* c = new forge_ctx();
  // set-up; a failure part-way releases what was made (forge_ctx_destroy
  // accepts a partly built context)
  const int rc = [&]() -> int {
    c->P = *p;
    c->pp = pair_params(*p);
    c->n = p->ndim;
    c->nnc = p->nnc;
    c->numk = p->numk;
    c->Nk = p->numk * p->numk;
    c->nL = p->nLower;
    c->nU = p->nUpper;
    c->nb = p->nUpper - p->nLower + 1;
    const i64 n = c->n, Nk = c->Nk, nnc = c->nnc;
    c->Coords.assign(Coords, Coords + 3 * n);
    c->tn.assign(tn, tn + 6);
    c->MomVal.assign(MomentaValues, MomentaValues + 2 * Nk);
    c->nMomComp.assign(nMomentaComponents, nMomentaComponents + 2 * Nk);
    c->u1.assign(nUnitCell_1, nUnitCell_1 + nnc);
    c->u2.assign(nUnitCell_2, nUnitCell_2 + nnc);
    c->opt.occ_mode = 0;
    c->opt.mix_mode = 0;
    c->opt.beta = 0.3;
    c->opt.etol = 0.0;
    FORGE_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    for (int q = 0; q < 16; ++q) FORGE_HIP(hipEventCreate(&c->ev[q]));
    const size_t nn = (size_t)n * n;
    CK(dalloc((void**)&c->X, sizeof(double) * 3 * n));
    FORGE_HIP(hipMemcpy(c->X, Coords, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    c->g.n = n;
    c->g.nnc = (int)nnc;
    c->g.X = c->X;
    c->g.Y = c->X + n;
    c->g.Z = c->X + 2 * n;
    for (int q = 0; q < MAXCELL; ++q) {
      c->g.cells.n1[q] = q < nnc ? (double)nUnitCell_1[q] : 0.0;
      c->g.cells.n2[q] = q < nnc ? (double)nUnitCell_2[q] : 0.0;
    }
    c->g.cells.t1x = tn[0];
    c->g.cells.t1y = tn[1];
    c->g.cells.t2x = tn[2];
    c->g.cells.t2y = tn[3];
    CK(dalloc((void**)&c->V, sizeof(double) * nnc * nn));
    CK(dalloc((void**)&c->T, sizeof(double) * nnc * nn));
    CK(pair_tables(c->g, c->pp, c->V, c->T, c->s));
    // layer bias, HamiltonianHartreeFock.f90:119-126
    std::vector<double> hb(n, 0.0);
    if (p->nlayers > 1) {
      i64 per = n / p->nlayers;
      for (i64 l = 0; l < p->nlayers; ++l)
        for (i64 i = 0; i < per; ++i)
          hb[i + l * per] = (double)((float)l - (float)(p->nlayers - 1) / 2.0f) * p->Delta /
                            (double)(float)(p->nlayers - 1);
    }
    CK(dalloc((void**)&c->bias, sizeof(double) * n));
    FORGE_HIP(hipMemcpy(c->bias, hb.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    CK(dalloc((void**)&c->pot, sizeof(double) * n));
    FORGE_HIP(hipMemset(c->pot, 0, sizeof(double) * n));
    CK(dalloc((void**)&c->rho, sizeof(z2) * nnc * nn));
    CK(dalloc((void**)&c->rho_new, sizeof(z2) * (nnc + 1) * nn));
    // k phases: assembly uses k.T_c (HartreeFock.f90:90), cells use the exact
    // grid phase -2 pi (i1 n1 + i2 (n1 + n2)) / numk (HartreeFock.f90:471)
    c->h_ph_asm.assign(Nk * nnc * 2, 0.0);
    c->h_ph_cell.assign(Nk * nnc * 2, 0.0);
    for (i64 k = 0; k < Nk; ++k) {
      i64 i1 = nMomentaComponents[k], i2 = nMomentaComponents[k + Nk];
      double vx = MomentaValues[i1 + c->numk * i2], vy = MomentaValues[i1 + c->numk * i2 + Nk];
      for (i64 cc = 0; cc < nnc; ++cc) {
        double n1 = (double)nUnitCell_1[cc], n2 = (double)nUnitCell_2[cc];
        double th = vx * (n1 * tn[0] + n2 * tn[2]) + vy * (n1 * tn[1] + n2 * tn[3]);
        c->h_ph_asm[(k * nnc + cc) * 2] = cos(th);
        c->h_ph_asm[(k * nnc + cc) * 2 + 1] = sin(th);
        double a = -2.0 * PI *
                   (double)(i1 * nUnitCell_1[cc] + i2 * nUnitCell_1[cc] + i2 * nUnitCell_2[cc]) /
                   (double)c->numk;
        c->h_ph_cell[(k * nnc + cc) * 2] = cos(a);
        c->h_ph_cell[(k * nnc + cc) * 2 + 1] = sin(a);
      }
    }
    CK(dalloc((void**)&c->ph_asm, sizeof(double) * Nk * nnc * 2));
    CK(dalloc((void**)&c->ph_cell, sizeof(double) * Nk * nnc * 2));
    FORGE_HIP(hipMemcpy(c->ph_asm, c->h_ph_asm.data(), sizeof(double) * Nk * nnc * 2, hipMemcpyHostToDevice));
    FORGE_HIP(hipMemcpy(c->ph_cell, c->h_ph_cell.data(), sizeof(double) * Nk * nnc * 2, hipMemcpyHostToDevice));
    // eigensolver batch size from free memory
    size_t fr = 0, tot = 0;
    FORGE_HIP(hipMemGetInfo(&fr, &tot));
    const i64 nU = c->nU;
    const bool two_stage = reduction_mode(0, n) == 2;
    size_t per_k = sizeof(z2) * pad_ld(n) * n + tridiag_workspace_bytes(n, 1) + dc_workspace_bytes(n, 1, nU) +
                   sizeof(double) * (n * nU + n) + sizeof(z2) * n * nU + backtr_workspace_bytes(n, 1, nU) +
                   (two_stage ? twostage_workspace_bytes(n, 1) : 0);
    size_t fixed_extra = sizeof(z2) * n * c->nb * Nk * 2 + ((size_t)64 << 20);
    size_t budget = fr > fixed_extra ? (size_t)((fr - fixed_extra) * 0.85) : 0;
    i64 Kb = (i64)(budget / per_k);
    if (Kb < 1) {
      set_error("forge_ctx_create: not enough device memory for one k-point");
      return -4;
    }
    c->Kb = (int)std::min<i64>(Kb, Nk);
    const int Kb_ = c->Kb;
    c->ldH = pad_ld(n);
    CK(dalloc((void**)&c->H, sizeof(z2) * c->ldH * n * Kb_));
    CK(dalloc(&c->trd_ws, tridiag_workspace_bytes(n, Kb_)));
    CK(tridiag_bind(c->trd, c->trd_ws, n, Kb_));
    c->trd.A = c->H;
    c->trd.sA = (size_t)c->ldH * n;
    c->trd.lda = c->ldH;
    if (two_stage) {
      CK(dalloc(&c->ts_ws, twostage_workspace_bytes(n, Kb_)));
      CK(twostage_bind(c->ts, c->ts_ws, n, Kb_));
    }
    CK(dalloc(&c->dc_ws, dc_workspace_bytes(n, Kb_, nU)));
    CK(dalloc((void**)&c->Zq, sizeof(double) * n * nU * Kb_));
    CK(dalloc((void**)&c->evals, sizeof(double) * n * Kb_));
    CK(dalloc((void**)&c->Cv, sizeof(z2) * n * nU * Kb_));
    CK(dalloc(&c->bt_ws, backtr_workspace_bytes(n, Kb_, nU)));
    FORGE_HIP(hipHostMalloc((void**)&c->hpin, sizeof(double) * (n * Kb_ + 2 * n + HPIN_SMALL),
                            hipHostMallocDefault));
    CK(dalloc((void**)&c->Cwin, sizeof(z2) * n * c->nb * Nk));
    CK(dalloc((void**)&c->Psi, sizeof(z2) * n * c->nb * Nk));
    CK(dalloc((void**)&c->wph, sizeof(double) * c->nb * Nk * nnc * 2));
    CK(dalloc((void**)&c->gidx, sizeof(int) * c->nb * Nk));
    CK(dalloc((void**)&c->part, sizeof(double) * 5 * 2048));
    CK(dalloc((void**)&c->e5, sizeof(double) * 16));
    CK(dalloc((void**)&c->dots, sizeof(double) * 16));
    CK(dalloc((void**)&c->da, sizeof(double) * n));
    CK(dalloc((void**)&c->db, sizeof(double) * n));
    CK(dalloc((void**)&c->ya, sizeof(double) * n));
    CK(dalloc((void**)&c->yb, sizeof(double) * n));
    FORGE_HIP(hipStreamSynchronize(c->s));
    return 0;
  }();
  if (rc != 0) {
    forge_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

extern "C" int forge_ctx_destroy(forge_ctx* c) {
  if (!c) return 0;
  // teardown (also of a partly built context): errors here leave nothing
  // for the caller to act on
  if (c->s) (void)hipStreamSynchronize(c->s);
  void* ptrs[] = {c->X, c->V, c->T, c->L, c->bias, c->pot, c->rho, c->rho_new, c->H, c->trd_ws, c->ts_ws,
                  c->dc_ws, c->Zq, c->evals, c->Cv, c->bt_ws, c->Cwin, c->ph_asm, c->ph_cell,
                  c->Psi, c->wph, c->gidx, c->part, c->e5, c->dots, c->da, c->db, c->ya, c->yb,
                  c->S, c->c3_perm, c->c3_tab, c->c3_ph_asm, c->c3_dexpand, c->c3_dpsi,
                  c->c3_Psi, c->c3_wph, c->c2_perm, c->valley_ws, c->dband, c->dS, c->docc,
                  c->dst, c->drow};
  for (auto& g : c->it_graph)
    if (g) (void)hipGraphExecDestroy(g);
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (c->hpin) (void)hipHostFree(c->hpin);
  tridiag_release(c->trd);
  twostage_release_events(c->ts);
  for (i64 q = 0; q < c->trd.nev; ++q) (void)hipEventDestroy(c->trd.evs[q]);
  delete[] c->trd.evs;
  for (int q = 0; q < 16; ++q)
    if (c->ev[q]) (void)hipEventDestroy(c->ev[q]);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
  return 0;
}

static int ensure_longrange(forge_ctx* c) {
  if (c->L) return 0;
  CK(dalloc((void**)&c->L, sizeof(double) * c->n * c->n));
  CK(longrange(c->g, c->pp, (int)c->P.numC, c->L, c->s));
  return 0;
}

extern "C" int forge_longrange(forge_ctx* c, double* LongRange) {
  if (!c || !LongRange) { set_error("forge_longrange: null argument"); return -2; }
  CK(ensure_longrange(c));
  FORGE_HIP(hipMemcpyAsync(LongRange, c->L, sizeof(double) * c->n * c->n, hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipStreamSynchronize(c->s));
  return 0;
}

static int hamiltonian_one(forge_ctx* c, forge_z* zH, const double* Potential, const forge_z* zFock,
                           const double* vk) {
  const i64 n = c->n, nnc = c->nnc;
  const size_t nn = (size_t)n * n;
  std::vector<double> ph(nnc * 2);
  for (i64 cc = 0; cc < nnc; ++cc) {
    double n1 = (double)c->u1[cc], n2 = (double)c->u2[cc];
    double th = vk[0] * (n1 * c->tn[0] + n2 * c->tn[2]) + vk[1] * (n1 * c->tn[1] + n2 * c->tn[3]);
    ph[cc * 2] = cos(th);
    ph[cc * 2 + 1] = sin(th);
  }
  double* dph = (double*)c->part;  // scratch
  FORGE_HIP(hipMemcpyAsync(dph, ph.data(), sizeof(double) * nnc * 2, hipMemcpyHostToDevice, c->s));
  // the potential goes to db and rho to rho_new: both are rebuilt by every
  // SCF iteration before they are read, so a Hamiltonian call between two
  // forge_scf_run calls leaves the SCF state (rho, da, pot, energies) intact
  double* vpot = c->db;
  if (Potential)
    FORGE_HIP(hipMemcpyAsync(vpot, Potential, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
  else
    FORGE_HIP(hipMemsetAsync(vpot, 0, sizeof(double) * n, c->s));
  if (zFock)
    FORGE_HIP(hipMemcpyAsync(c->rho_new, zFock, sizeof(z2) * nnc * nn, hipMemcpyHostToDevice, c->s));
  CK(assemble(n, (int)nnc, 1, zFock ? c->rho_new : nullptr, c->V, c->T, vpot, c->bias,
              c->P.alpha, dph, c->H, c->ldH, (size_t)c->ldH * n, c->s));
  hipLaunchKernelGGL(k_zero_upper, dim3(cdiv(n * n, 256)), dim3(256), 0, c->s, n, c->ldH, c->H);
  FORGE_LAUNCH();
  FORGE_HIP(hipMemcpy2DAsync(zH, sizeof(z2) * n, c->H, sizeof(z2) * c->ldH, sizeof(z2) * n, n,
                             hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipStreamSynchronize(c->s));
  return 0;
}

extern "C" int forge_tb_hamiltonian(forge_ctx* c, forge_z* zH, const double* vk) {
  if (!c || !zH || !vk) { set_error("forge_tb_hamiltonian: null argument"); return -2; }
  return hamiltonian_one(c, zH, nullptr, nullptr, vk);
}

extern "C" int forge_hamiltonian_hf(forge_ctx* c, forge_z* zH, const double* Potential,
                                    const forge_z* zFock, const double* vk) {
  if (!c || !zH || !Potential || !zFock || !vk) {
    set_error("forge_hamiltonian_hf: null argument");
    return -2;
  }
  return hamiltonian_one(c, zH, Potential, zFock, vk);
}

// Eigenpairs 1..mout of the Kb' matrices in c->H (lower): evals (Kb x n),
// Cv (Kb x n x nU, ld n).  Returns D&C failures (> 0) or < 0.
// precision 1 before the switch: FP32 copy in the hemv and FP32 back-transform
static bool mixed_phase(const forge_ctx* c) {
  return c->opt.precision == 1 && !c->mixed_done && c->S;
}

static int eig_batch(forge_ctx* c, int K, i64 mout, bool timed) {
  const i64 n = c->n;
  c->trd.K = K;
  c->trd.ts = (reduction_mode(c->opt.reduction, n) == 2 && c->ts_ws) ? &c->ts : nullptr;
  const bool mixed = mixed_phase(c);
  c->trd.S = mixed ? c->S : nullptr;
  c->trd.use_graph = c->opt.graph == 1;
  c->trd.sS = (size_t)c->ldH * n;
  double hemv_ms = 0, hemv_bytes = 0;
  long hemv_n = 0;
  if (timed) FORGE_HIP(hipEventRecord(c->ev[2], c->s));
  if (c->dev_iter) CK(tridiag_enqueue(c->trd, c->s));
  else
    CK(tridiag(c->trd, c->s, timed ? &hemv_ms : nullptr, timed ? &hemv_n : nullptr,
               timed ? &hemv_bytes : nullptr));
  if (timed) FORGE_HIP(hipEventRecord(c->ev[3], c->s));
  // device iteration: the failure count stays on the device (read after it)
  int fails = dc_solve(c->trd.d, c->trd.e, n, K, mout, c->evals, c->Zq, c->dc_ws, c->s, c->dev_iter);
  if (fails < 0) return fails;
  if (timed) FORGE_HIP(hipEventRecord(c->ev[4], c->s));
  // mixed phase: the FP32 copy is free once the reduction is done and holds
  // the FP32 eigenvectors (n x mout x K <= ldH x n x K)
  int width = 0;
  if (c->trd.ts) c->trd.ts->q2_timed = timed && c->trd.two_stage;
  const int btr = backtransform(c->trd, c->Zq, n, (size_t)n * mout, mout, c->Cv, n,
                                (size_t)n * c->nU, c->bt_ws, c->s, mixed ? c->S : nullptr, &width);
  if (c->trd.ts) c->trd.ts->q2_timed = false;
  CK(btr);
  c->tm[13] = width;
  {
    // the hemv variants of the last reduction, one decimal digit per band of
    // trailing orders (lowest band last)
    int code[4];
    i64 lo[4];
    (void)tridiag_hemv_plan(c->trd, code, lo);
    c->tm[14] = 1000.0 * code[3] + 100.0 * code[2] + 10.0 * code[1] + code[0];
  }
  if (timed) {
    FORGE_HIP(hipEventRecord(c->ev[5], c->s));
    FORGE_HIP(hipEventSynchronize(c->ev[5]));
    float a, b, d;
    FORGE_HIP(hipEventElapsedTime(&a, c->ev[2], c->ev[3]));
    FORGE_HIP(hipEventElapsedTime(&b, c->ev[3], c->ev[4]));
    FORGE_HIP(hipEventElapsedTime(&d, c->ev[4], c->ev[5]));
    c->tm[1] += a;
    c->tm[2] += b;
    c->tm[3] += d;
    c->tm[9] += hemv_n;
    c->tm[10] += hemv_ms;
    c->tm[11] += hemv_bytes;
    if (c->trd.ts && c->trd.two_stage) {
      long q2n = 0;
      double q2ms = 0.0, q2fl = 0.0, q2b = 0.0;
      CK(twostage_q2_timing(*c->trd.ts, &q2n, &q2ms, &q2fl, &q2b));
      c->tm[15] += q2n;
      c->tm[16] += q2ms;
      c->tm[17] += q2fl;
      c->tm[18] += q2b;
    }
  }
  return fails;
}

extern "C" int forge_zdiagred(forge_z* A, int64_t n, double* eigenvalues, int64_t il, int64_t iu,
                              int32_t* info) {
  if (!A || !eigenvalues || n < 1 || il < 1 || iu < il || iu > n) {
    set_error("forge_zdiagred: invalid argument");
    return -2;
  }
  if (forge_device_count() < 1) { set_error("forge_zdiagred: no HIP device"); return -3; }
  // a stand-alone one-matrix solve; every buffer is released on every path
  hipStream_t s = nullptr;
  FORGE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  z2* H = nullptr; void* tw = nullptr; void* dw = nullptr; double* Zq = nullptr; double* ev = nullptr;
  z2* Cv = nullptr; void* bw = nullptr; void* tsw = nullptr;
  TridiagWork t{};
  TwoStageWork ts;
  const int rc = [&]() -> int {
    const i64 ld = pad_ld(n);
    CK(dalloc((void**)&H, sizeof(z2) * ld * n));
    CK(dalloc(&tw, tridiag_workspace_bytes(n, 1)));
    if (reduction_mode(0, n) == 2) {
      CK(dalloc(&tsw, twostage_workspace_bytes(n, 1)));
      CK(twostage_bind(ts, tsw, n, 1));
    }
    CK(dalloc(&dw, dc_workspace_bytes(n, 1, iu)));
    CK(dalloc((void**)&Zq, sizeof(double) * n * iu));
    CK(dalloc((void**)&ev, sizeof(double) * n));
    CK(dalloc((void**)&Cv, sizeof(z2) * n * iu));
    CK(dalloc(&bw, backtr_workspace_bytes(n, 1, iu)));
    CK(tridiag_bind(t, tw, n, 1));
    t.A = H;
    t.sA = (size_t)ld * n;
    t.lda = ld;
    if (tsw) t.ts = &ts;
    FORGE_HIP(hipMemcpy2DAsync(H, sizeof(z2) * ld, A, sizeof(z2) * n, sizeof(z2) * n, n,
                               hipMemcpyHostToDevice, s));
    CK(tridiag(t, s, nullptr, nullptr, nullptr));
    const int fails = dc_solve(t.d, t.e, n, 1, iu, ev, Zq, dw, s);
    if (fails < 0) return fails;
    CK(backtransform(t, Zq, n, (size_t)n * iu, iu, Cv, n, (size_t)n * iu, bw, s));
    const i64 m = iu - il + 1;
    FORGE_HIP(hipMemcpyAsync(A, Cv + n * (il - 1), sizeof(z2) * n * m, hipMemcpyDeviceToHost, s));
    std::vector<double> hv(n);
    FORGE_HIP(hipMemcpyAsync(hv.data(), ev, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    FORGE_HIP(hipStreamSynchronize(s));
    for (i64 q = 0; q < m; ++q) eigenvalues[q] = hv[il - 1 + q];
    if (info) *info = fails;
    return fails;
  }();
  // teardown: errors here leave nothing for the caller to act on
  (void)hipStreamSynchronize(s);
  for (void* q : {(void*)H, tw, dw, (void*)Zq, (void*)ev, (void*)Cv, bw, tsw})
    if (q) (void)hipFree(q);
  tridiag_release(t);
  for (i64 q = 0; q < t.nev; ++q) (void)hipEventDestroy(t.evs[q]);
  delete[] t.evs;
  (void)hipStreamDestroy(s);
  return rc;
}

// ---------------------------------------------------------------- products of the LAPACK shim
// The reference's product routines (LapackRoutines.f90:119-236) on the
// matrix-core GEMMs of gemm.hip, for host arrays: each call copies its
// operands to the device and the result back (stateless, like zdiagRed).

// C += z R (or C = z R when overwrite), n elements
__global__ void k_zaxpy(size_t n, z2 zf, const z2* __restrict__ R, z2* __restrict__ C, int overwrite) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  const z2 v = zf * R[q];
  C[q] = overwrite ? v : C[q] + v;
}

namespace {
// device buffers and a stream of one host-array call, released on every path
struct Scoped {
  hipStream_t s = nullptr;
  std::vector<void*> bufs;
  ~Scoped() {
    if (s) (void)hipStreamSynchronize(s);
    for (void* b : bufs) (void)hipFree(b);
    if (s) (void)hipStreamDestroy(s);
  }
  int init() {
    FORGE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return 0;
  }
  template <typename T> int alloc(T** p, size_t count) {
    void* q = nullptr;
    FORGE_HIP(hipMalloc(&q, sizeof(T) * (count ? count : 1)));
    bufs.push_back(q);
    *p = (T*)q;
    return 0;
  }
};
}  // namespace

extern "C" int forge_z_vector_mul(forge_z* C, const forge_z* A, int64_t n, int64_t k,
                                  const forge_z* z_factor) {
  if (!C || !A || n < 1 || k < 1) { set_error("forge_z_vector_mul: invalid argument"); return -2; }
  if (forge_device_count() < 1) { set_error("forge_z_vector_mul: no HIP device"); return -3; }
  Scoped sc;
  CK(sc.init());
  z2 *dA, *dR, *dC;
  const size_t nn = (size_t)n * n;
  CK(sc.alloc(&dA, (size_t)n * k));
  CK(sc.alloc(&dR, nn));
  CK(sc.alloc(&dC, nn));
  FORGE_HIP(hipMemcpyAsync(dA, A, sizeof(z2) * n * k, hipMemcpyHostToDevice, sc.s));
  FORGE_HIP(hipMemcpyAsync(dC, C, sizeof(z2) * nn, hipMemcpyHostToDevice, sc.s));
  CK(zgemm(Op::N, Op::H, n, n, k, 1.0, dA, n, 0, dA, n, 0, 0.0, dR, n, 0, 1, sc.s));
  const z2 zf = z_factor ? mk(z_factor->re, z_factor->im) : mk(1.0, 0.0);
  hipLaunchKernelGGL(k_zaxpy, dim3((unsigned)cdiv((i64)nn, 256)), dim3(256), 0, sc.s, nn, zf, dR, dC, 0);
  FORGE_LAUNCH();
  FORGE_HIP(hipMemcpyAsync(C, dC, sizeof(z2) * nn, hipMemcpyDeviceToHost, sc.s));
  FORGE_HIP(hipStreamSynchronize(sc.s));
  return 0;
}

extern "C" int forge_z_matrix_mul(forge_z* C, const forge_z* A, const forge_z* B, int64_t m,
                                  int64_t k, int64_t n) {
  if (!C || !A || !B || m < 1 || k < 1 || n < 1) {
    set_error("forge_z_matrix_mul: invalid argument");
    return -2;
  }
  if (forge_device_count() < 1) { set_error("forge_z_matrix_mul: no HIP device"); return -3; }
  Scoped sc;
  CK(sc.init());
  z2 *dA, *dB, *dC;
  CK(sc.alloc(&dA, (size_t)m * k));
  CK(sc.alloc(&dB, (size_t)k * n));
  CK(sc.alloc(&dC, (size_t)m * n));
  FORGE_HIP(hipMemcpyAsync(dA, A, sizeof(z2) * m * k, hipMemcpyHostToDevice, sc.s));
  FORGE_HIP(hipMemcpyAsync(dB, B, sizeof(z2) * k * n, hipMemcpyHostToDevice, sc.s));
  CK(zgemm(Op::N, Op::N, m, n, k, 1.0, dA, m, 0, dB, k, 0, 0.0, dC, m, 0, 1, sc.s));
  FORGE_HIP(hipMemcpyAsync(C, dC, sizeof(z2) * m * n, hipMemcpyDeviceToHost, sc.s));
  FORGE_HIP(hipStreamSynchronize(sc.s));
  return 0;
}

extern "C" int forge_z_matrix_overlap(forge_z* C, const forge_z* A, const forge_z* B, int64_t m1,
                                      int64_t m0, const forge_z* z_factor) {
  if (!C || !A || !B || m1 < 1 || m0 < 1) {
    set_error("forge_z_matrix_overlap: invalid argument");
    return -2;
  }
  if (forge_device_count() < 1) { set_error("forge_z_matrix_overlap: no HIP device"); return -3; }
  Scoped sc;
  CK(sc.init());
  z2 *dA, *dB, *dR;
  const size_t mm = (size_t)m0 * m0;
  CK(sc.alloc(&dA, (size_t)m1 * m0));
  CK(sc.alloc(&dB, (size_t)m1 * m0));
  CK(sc.alloc(&dR, 2 * mm));
  FORGE_HIP(hipMemcpyAsync(dA, A, sizeof(z2) * m1 * m0, hipMemcpyHostToDevice, sc.s));
  FORGE_HIP(hipMemcpyAsync(dB, B, sizeof(z2) * m1 * m0, hipMemcpyHostToDevice, sc.s));
  CK(zgemm(Op::H, Op::N, m0, m0, m1, 1.0, dA, m1, 0, dB, m1, 0, 0.0, dR, m0, 0, 1, sc.s));
  const z2 zf = z_factor ? mk(z_factor->re, z_factor->im) : mk(1.0, 0.0);
  hipLaunchKernelGGL(k_zaxpy, dim3((unsigned)cdiv((i64)mm, 256)), dim3(256), 0, sc.s, mm, zf, dR,
                     dR + mm, 1);
  FORGE_LAUNCH();
  FORGE_HIP(hipMemcpyAsync(C, dR + mm, sizeof(z2) * mm, hipMemcpyDeviceToHost, sc.s));
  FORGE_HIP(hipStreamSynchronize(sc.s));
  return 0;
}

extern "C" int forge_c_matrix_mul(forge_c* C, const forge_c* A, const forge_c* B, int64_t m,
                                  int64_t k, int64_t n) {
  if (!C || !A || !B || m < 1 || k < 1 || n < 1) {
    set_error("forge_c_matrix_mul: invalid argument");
    return -2;
  }
  if (forge_device_count() < 1) { set_error("forge_c_matrix_mul: no HIP device"); return -3; }
  Scoped sc;
  CK(sc.init());
  float2 *dA, *dB, *dC;
  CK(sc.alloc(&dA, (size_t)m * k));
  CK(sc.alloc(&dB, (size_t)k * n));
  CK(sc.alloc(&dC, (size_t)m * n));
  FORGE_HIP(hipMemcpyAsync(dA, A, sizeof(float2) * m * k, hipMemcpyHostToDevice, sc.s));
  FORGE_HIP(hipMemcpyAsync(dB, B, sizeof(float2) * k * n, hipMemcpyHostToDevice, sc.s));
  CK(cgemm(Op::N, Op::N, m, n, k, 1.0, dA, m, 0, dB, k, 0, 0.0, dC, m, 0, 1, sc.s));
  FORGE_HIP(hipMemcpyAsync(C, dC, sizeof(float2) * m * n, hipMemcpyDeviceToHost, sc.s));
  FORGE_HIP(hipStreamSynchronize(sc.s));
  return 0;
}

// ---------------------------------------------------------------- C3-reduced k-set
// Tables of Solve_C3 / GetFock_C3 (HartreeFock.f90:131-398, 564-881) for one
// RotateLayers / nC3pairs (1-based, as C3_RelatedPoints returns it), rebuilt
// only when those change.  Corner factor sets per flat momentum k (6 each):
//   6k+0  f: first corner e^{i rot t1}, t1 = 2 pi (i1+i2)/numk (C3 block image)
//   6k+1  u = f g, g: first corner e^{-i rot t2}, t2 = 2 pi i1/numk (C3^2 image)
//   6k+2, 6k+3  conj(f), conj(u): the images of a state vector
//   6k+4  zphase of the C3^2 partner's eigenvectors, e^{i rot 2 pi (j1+j2)/numk}
//   6k+5  zphase of the C3 partner's eigenvectors, e^{i rot 2 pi j2/numk}
// ((j1, j2) the partner's components, fphase = 0); second corners conjugate.
static int c3_configure(forge_ctx* c, const int64_t* rot, const int64_t* pairs, const char* who) {
  const i64 n = c->n, Nk = c->Nk, numk = c->numk, nnc = c->nnc, nb = c->nb;
  const i64 nl = c->P.nlayers;
  if (!rot || !pairs) { set_error(std::string(who) + ": null RotateLayers or nC3pairs"); return -2; }
  if (numk % 3 != 0) {
    set_error(std::string(who) + ": numk must be a multiple of 3 (the C3-invariant momenta are "
              "the flat indices of Gamma, K and K')");
    return -2;
  }
  if (nl < 1 || nl > 64 || n % nl != 0 || (n / nl) % 2 != 0) {
    set_error(std::string(who) + ": layers must split ndim into two equal sublattice halves");
    return -2;
  }
  for (i64 k = 0; k < Nk; ++k)
    if (c->nMomComp[k] * numk + c->nMomComp[k + Nk] != k) {
      set_error(std::string(who) + ": momenta must be in SampleBZ order (flat index i1*numk + i2)");
      return -2;
    }
  std::vector<int64_t> key(rot, rot + nl);
  key.insert(key.end(), pairs, pairs + n);
  if (c->c3_ready && key == c->c3_key) return 0;
  const i64 per = n / nl, half = per / 2;
  std::vector<int> p(2 * n);
  std::vector<char> seen(n, 0);
  for (i64 i = 0; i < n; ++i) {
    const int64_t j = pairs[i] - 1;
    if (j < 0 || j >= n || seen[j]) {
      set_error(std::string(who) + ": nC3pairs is not a permutation of 1..ndim");
      return -2;
    }
    seen[j] = 1;
    p[i] = (int)j;
  }
  for (i64 l = 0; l < nl; ++l)
    if (p[l * per + half - 1] != l * per + half - 1 || p[l * per + per - 1] != l * per + per - 1) {
      set_error(std::string(who) + ": nC3pairs must fix the two corner sites of each layer");
      return -2;
    }
  for (i64 i = 0; i < n; ++i) p[n + i] = p[p[i]];
  // orbits, HartreeFock.f90:152-189
  c->c3_k2.assign(Nk, 0);
  c->c3_k3.assign(Nk, 0);
  c->c3_inv.assign(Nk, 0);
  c->c3_reps.clear();
  std::vector<char> incl(Nk, 0);
  for (i64 k = 0; k < Nk; ++k) {
    const i64 i1 = c->nMomComp[k], i2 = c->nMomComp[k + Nk];
    const i64 m = ((-i1 - i2) % numk + numk) % numk;
    c->c3_k2[k] = m * numk + i1 % numk;
    c->c3_k3[k] = (i2 % numk) * numk + m;
    c->c3_inv[k] = k == 0 || k == numk / 3 * (numk + 1) || k == 2 * numk / 3 * (numk + 1);
  }
  for (i64 k = 0; k < Nk; ++k) {
    if (incl[k]) continue;
    c->c3_reps.push_back(k);
    incl[k] = incl[c->c3_k2[k]] = incl[c->c3_k3[k]] = 1;
  }
  const i64 R = (i64)c->c3_reps.size();
  // corner factors
  std::vector<z2> tab((size_t)6 * Nk * nl * 2);
  auto put = [&](i64 set, double a) {
    for (i64 l = 0; l < nl; ++l) {
      const double r = (rot[l] == 1 || rot[l] == -1) ? (double)rot[l] : 0.0;
      tab[(set * nl + l) * 2] = r != 0.0 ? mk(std::cos(r * a), std::sin(r * a)) : mk(1.0, 0.0);
      tab[(set * nl + l) * 2 + 1] = r != 0.0 ? mk(std::cos(r * a), -std::sin(r * a)) : mk(1.0, 0.0);
    }
  };
  for (i64 k = 0; k < Nk; ++k) {
    const i64 i1 = c->nMomComp[k], i2 = c->nMomComp[k + Nk];
    const i64 k2 = c->c3_k2[k], k3 = c->c3_k3[k];
    const double t1 = 2.0 * PI * (double)(i1 + i2) / (double)numk;
    const double t2 = 2.0 * PI * (double)i1 / (double)numk;
    put(6 * k + 0, t1);
    put(6 * k + 1, t1 - t2);
    put(6 * k + 2, -t1);
    put(6 * k + 3, -(t1 - t2));
    put(6 * k + 4, 2.0 * PI * (double)(c->nMomComp[k2] + c->nMomComp[k2 + Nk]) / (double)numk);
    put(6 * k + 5, 2.0 * PI * (double)c->nMomComp[k3 + Nk] / (double)numk);
  }
  // partners' window columns and the per-representative bulk slots
  c->c3_expand.clear();
  c->c3_slots.clear();
  for (i64 r = 0; r < R; ++r) {
    const i64 k = c->c3_reps[r];
    const i64 k2 = c->c3_k2[k], k3 = c->c3_k3[k];
    C3Slot sl;
    sl.ph0 = nullptr;  // device pointers set below
    sl.fset = (int)(6 * k + 0);
    sl.uset = (int)(6 * k + 1);
    sl.w = c->c3_inv[k] ? 1.0 / 3.0 : 1.0;
    c->c3_slots.push_back(sl);
    if (c->c3_inv[k]) continue;
    for (i64 b = 0; b < nb; ++b) {
      c->c3_expand.push_back(C3Col{b + nb * k, b + nb * k2, 2, (int)(6 * k + 4)});
      c->c3_expand.push_back(C3Col{b + nb * k, b + nb * k3, 1, (int)(6 * k + 5)});
    }
  }
  // device tables (allocated once per context; sizes depend only on n, Nk)
  if (!c->c3_perm) {
    c->c3_cap = nb * (Nk + 6);
    CK(dalloc((void**)&c->c3_perm, sizeof(int) * 2 * n));
    CK(dalloc((void**)&c->c3_tab, sizeof(z2) * tab.size()));
    CK(dalloc((void**)&c->c3_ph_asm, sizeof(double) * Nk * nnc * 2));
    CK(dalloc((void**)&c->c3_dexpand, sizeof(C3Col) * std::max<i64>(2 * nb * Nk, 1)));
    CK(dalloc((void**)&c->c3_dpsi, sizeof(C3Col) * c->c3_cap));
    CK(dalloc((void**)&c->c3_Psi, sizeof(z2) * n * c->c3_cap));
    CK(dalloc((void**)&c->c3_wph, sizeof(double) * c->c3_cap * nnc * 2));
  }
  std::vector<double> pa((size_t)R * nnc * 2);
  for (i64 r = 0; r < R; ++r)
    std::memcpy(&pa[(size_t)r * nnc * 2], &c->h_ph_asm[(size_t)c->c3_reps[r] * nnc * 2],
                sizeof(double) * nnc * 2);
  for (i64 r = 0; r < R; ++r) {
    const i64 k = c->c3_reps[r];
    c->c3_slots[r].ph0 = c->ph_cell + k * nnc * 2;
    c->c3_slots[r].ph1 = c->ph_cell + c->c3_k3[k] * nnc * 2;
    c->c3_slots[r].ph2 = c->ph_cell + c->c3_k2[k] * nnc * 2;
  }
  FORGE_HIP(hipMemcpy(c->c3_perm, p.data(), sizeof(int) * 2 * n, hipMemcpyHostToDevice));
  FORGE_HIP(hipMemcpy(c->c3_tab, tab.data(), sizeof(z2) * tab.size(), hipMemcpyHostToDevice));
  FORGE_HIP(hipMemcpy(c->c3_ph_asm, pa.data(), sizeof(double) * pa.size(), hipMemcpyHostToDevice));
  if (!c->c3_expand.empty())
    FORGE_HIP(hipMemcpy(c->c3_dexpand, c->c3_expand.data(), sizeof(C3Col) * c->c3_expand.size(),
                        hipMemcpyHostToDevice));
  c->c3g = C3Geom{n, per, (int)nl, c->c3_perm, c->c3_tab};
  c->c3_key = key;
  c->c3_ready = true;
  return 0;
}

// Solve_C3 on the device: assembly, eigensolver and bulk density for the
// orbit representatives only, then the partners' window columns as images.
static int solve_device_c3(forge_ctx* c, const z2* rho_dev, const double* pot_dev,
                           std::vector<double>& hband, bool timed) {
  const i64 n = c->n, nnc = c->nnc, Nk = c->Nk, nU = c->nU, nb = c->nb;
  const i64 R = (i64)c->c3_reps.size();
  hband.assign(nb * Nk, 0.0);
  double* hev = c->hpin;  // n x Kb
  int fails = 0;
  for (i64 r0 = 0; r0 < R; r0 += c->Kb) {
    const int K = (int)std::min<i64>(c->Kb, R - r0);
    if (timed) FORGE_HIP(hipEventRecord(c->ev[0], c->s));
    CK(assemble(n, (int)nnc, K, rho_dev, c->V, c->T, pot_dev, c->bias, c->P.alpha,
                c->c3_ph_asm + r0 * nnc * 2, c->H, c->ldH, (size_t)c->ldH * n, c->s));
    if (timed) {
      FORGE_HIP(hipEventRecord(c->ev[1], c->s));
      FORGE_HIP(hipEventSynchronize(c->ev[1]));
      float a;
      FORGE_HIP(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
      c->tm[0] += a;
      c->tm[7] += 1;
      c->tm[8] += a;
    }
    int f = eig_batch(c, K, nU, timed);
    if (f < 0) return f;
    fails += f;
    if (timed) FORGE_HIP(hipEventRecord(c->ev[6], c->s));
    CK(bulk_density_c3(c->c3g, K, c->Cv, n, (size_t)n * nU, c->nL - 1, (int)nnc,
                       c->c3_slots.data() + r0, c->rho_new, r0 > 0, c->s));
    for (int kb = 0; kb < K; ++kb) {
      hipLaunchKernelGGL(k_copy_window, dim3(cdiv(n * nb, 256), 1), dim3(256), 0, c->s, n, nb,
                         c->nL - 1, 1, c->c3_reps[r0 + kb], c->Cv + (size_t)kb * n * nU, (size_t)n,
                         (size_t)n * nU, c->Cwin);
      FORGE_LAUNCH();
    }
    if (timed) {
      FORGE_HIP(hipEventRecord(c->ev[7], c->s));
      FORGE_HIP(hipEventSynchronize(c->ev[7]));
      float a;
      FORGE_HIP(hipEventElapsedTime(&a, c->ev[6], c->ev[7]));
      c->tm[4] += a;
    }
    FORGE_HIP(hipMemcpyAsync(hev, c->evals, sizeof(double) * n * K, hipMemcpyDeviceToHost, c->s));
    FORGE_HIP(hipStreamSynchronize(c->s));
    for (int kb = 0; kb < K; ++kb) {
      const i64 k = c->c3_reps[r0 + kb];
      for (i64 b = 0; b < nb; ++b) {
        const double e = hev[kb * n + (c->nL - 1) + b];
        hband[b + nb * k] = e;
        hband[b + nb * c->c3_k2[k]] = e;
        hband[b + nb * c->c3_k3[k]] = e;
      }
    }
  }
  // partners' eigenvectors (HartreeFock.f90:212-277): sources are the
  // representatives' columns, destinations the partners' (disjoint)
  CK(c3_columns(c->c3g, (i64)c->c3_expand.size(), c->c3_dexpand, c->Cwin, c->Cwin, c->s));
  return fails;
}

// Solve on the device: rho_dev (nnc cells) and pot_dev in, bulk density
// (unnormalised, all cells) into c->rho_new, window bands to hband (nb x Nk,
// host) and window eigenvectors to c->Cwin.
// eig_ready: a sweep has already assembled and solved all Nk matrices (one
// batch) into c->evals / c->Cv; only the densities and bands are formed here.
static int solve_device(forge_ctx* c, const z2* rho_dev, const double* pot_dev,
                        std::vector<double>& hband, bool timed, bool eig_ready = false) {
  const i64 n = c->n, nnc = c->nnc, Nk = c->Nk, nU = c->nU, nb = c->nb;
  const size_t nn = (size_t)n * n;
  hband.assign(nb * Nk, 0.0);
  double* hev = c->hpin;  // n x Kb
  int fails = 0;
  for (i64 k0 = 0; k0 < Nk; k0 += c->Kb) {
    const int K = (int)std::min<i64>(c->Kb, Nk - k0);
    if (!eig_ready) {
      if (timed) FORGE_HIP(hipEventRecord(c->ev[0], c->s));
      CK(assemble(n, (int)nnc, K, rho_dev, c->V, c->T, pot_dev, c->bias, c->P.alpha,
                  c->ph_asm + k0 * nnc * 2, c->H, c->ldH, (size_t)c->ldH * n, c->s));
      if (timed) {
        FORGE_HIP(hipEventRecord(c->ev[1], c->s));
        FORGE_HIP(hipEventSynchronize(c->ev[1]));
        float a;
        FORGE_HIP(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
        c->tm[0] += a;
        c->tm[7] += 1;
        c->tm[8] += a;
      }
      int f = eig_batch(c, K, nU, timed);
      if (f < 0) return f;
      fails += f;
    }
    if (timed) FORGE_HIP(hipEventRecord(c->ev[6], c->s));
    // bulk density: bands 1..nLower-1, phases of k0..k0+K-1, summed over k
    // (the first k of the first batch overwrites)
    {
      // bulk_density overwrites for its first k; for later batches accumulate
      CK(bulk_density(n, K, c->Cv, n, (size_t)n * nU, c->nL - 1, (int)nnc,
                      c->ph_cell + k0 * nnc * 2, c->rho_new, k0 > 0, c->s));
    }
    hipLaunchKernelGGL(k_copy_window, dim3(cdiv(n * nb, 256), K), dim3(256), 0, c->s, n, nb,
                       c->nL - 1, K, k0, c->Cv, (size_t)n, (size_t)n * nU, c->Cwin);
    FORGE_LAUNCH();
    if (timed) {
      FORGE_HIP(hipEventRecord(c->ev[7], c->s));
      FORGE_HIP(hipEventSynchronize(c->ev[7]));
      float a;
      FORGE_HIP(hipEventElapsedTime(&a, c->ev[6], c->ev[7]));
      c->tm[4] += a;
    }
    if (c->dev_iter) {
      // the window energies stay on the device for k_occupations
      CK(band_gather(n, nb, c->nL - 1, K, k0, Nk, c->evals, c->dband, c->s));
      continue;
    }
    FORGE_HIP(hipMemcpyAsync(hev, c->evals, sizeof(double) * n * K, hipMemcpyDeviceToHost, c->s));
    FORGE_HIP(hipStreamSynchronize(c->s));
    for (int kb = 0; kb < K; ++kb)
      for (i64 b = 0; b < nb; ++b) hband[b + nb * (k0 + kb)] = hev[kb * n + (c->nL - 1) + b];
  }
  return fails;
}

extern "C" int forge_solve(forge_ctx* c, forge_z* zFockBulk, const forge_z* zFock,
                           const double* Potential, double* Bands, forge_z* zEigenvectors) {
  if (!c || !zFockBulk || !zFock || !Potential || !Bands || !zEigenvectors) {
    set_error("forge_solve: null argument");
    return -2;
  }
  const i64 n = c->n, nnc = c->nnc, Nk = c->Nk, nb = c->nb;
  const size_t nn = (size_t)n * n;
  FORGE_HIP(hipMemcpyAsync(c->rho, zFock, sizeof(z2) * nnc * nn, hipMemcpyHostToDevice, c->s));
  FORGE_HIP(hipMemcpyAsync(c->db, Potential, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
  std::vector<double> hb;
  int f = solve_device(c, c->rho, c->db, hb, false);
  if (f < 0) return f;
  FORGE_HIP(hipMemcpyAsync(zFockBulk, c->rho_new, sizeof(z2) * nnc * nn, hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipMemcpyAsync(zEigenvectors, c->Cwin, sizeof(z2) * n * nb * Nk, hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipStreamSynchronize(c->s));
  std::memcpy(Bands, hb.data(), sizeof(double) * nb * Nk);
  c->scf_ready = false;  // device rho buffers were reused
  return f;
}

// ---------------------------------------------------------------- host-side sort / Fermi
// dsort_lists, TightBinding.f90:984-1027 (heap sort; the value column carries
// the flat index (b-1)*Nk + k).
static void heap_sort2(std::vector<double>& key, std::vector<double>& val) {
  const i64 N = (i64)key.size();
  if (N < 2) return;
  i64 l = N / 2 + 1, ir = N, i, j;
  double ck, cv;
  for (;;) {
    if (l > 1) {
      --l; ck = key[l - 1]; cv = val[l - 1];
    } else {
      ck = key[ir - 1]; cv = val[ir - 1];
      key[ir - 1] = key[0]; val[ir - 1] = val[0];
      if (--ir == 1) { key[0] = ck; val[0] = cv; return; }
    }
    i = l; j = 2 * l;
    while (j <= ir) {
      if (j < ir && key[j - 1] < key[j]) ++j;
      if (ck < key[j - 1]) { key[i - 1] = key[j - 1]; val[i - 1] = val[j - 1]; i = j; j = 2 * j; }
      else j = ir + 1;
    }
    key[i - 1] = ck; val[i - 1] = cv;
  }
}

struct Occupation {
  i64 nF = 0, nOcc = 0, nPart = 0;
  double EF = 0, Deg = 0;
  std::vector<int> idx;       // Cwin column of each weighted state
  std::vector<double> wph;    // [S][nnc][2]
};

// SortEnergies_1Spin (HartreeFock.f90:885-925): the window energies in
// ascending order (key) and the 1-based flat source of each (val).
static void sort_window(const forge_ctx* c, const std::vector<double>& band, std::vector<double>& key,
                        std::vector<double>& val) {
  const i64 Nk = c->Nk, nb = c->nb, N = nb * Nk;
  key.resize(N);
  val.resize(N);
  for (i64 ic = 1; ic <= N; ++ic) {
    i64 indev = (ic - 1) / Nk + 1;
    key[ic - 1] = band[(indev - 1) + nb * (ic - (indev - 1) * Nk - 1)];
    val[ic - 1] = (double)ic;
  }
  heap_sort2(key, val);
}

// The weighted states of GetFock (HartreeFock.f90:506-555): for each sorted
// position q with a weight w[q] + i wi[q] != 0, its Cwin column and its nnc
// complex weights (weight x cell phase of its momentum).
static void weighted_states(const forge_ctx* c, const std::vector<double>& val,
                            const std::vector<double>& w, const std::vector<double>& wi,
                            Occupation& o) {
  const i64 Nk = c->Nk, nb = c->nb, N = nb * Nk, nnc = c->nnc;
  o.idx.clear();
  o.wph.clear();
  for (i64 q = 0; q < N; ++q) {
    if (w[q] == 0.0 && wi[q] == 0.0) continue;
    i64 src = (i64)std::llround(val[q]);
    i64 indev = (src - 1) / Nk + 1, k = src - (indev - 1) * Nk;  // band indev, k-point k (1-based)
    o.idx.push_back((int)((indev - 1) + nb * (k - 1)));
    i64 i1 = c->nMomComp[k - 1], i2 = c->nMomComp[k - 1 + Nk];
    for (i64 cc = 0; cc < nnc; ++cc) {
      double a = -2.0 * PI * (double)(i1 * c->u1[cc] + i2 * c->u1[cc] + i2 * c->u2[cc]) / (double)c->numk;
      std::complex<double> z = std::complex<double>(std::cos(a), std::sin(a)) * std::complex<double>(w[q], wi[q]);
      o.wph.push_back(z.real());
      o.wph.push_back(z.imag());
    }
  }
}

// T = 0 weights of the first nOcc sorted states (GetFock,
// HartreeFock.f90:506-555): 1, and DegFactor for the last nPart of them
// (occ_mode 0: the reference's zdeg = (float(DegFactor), 8); 1: real).
static void zero_temperature_weights(int occ_mode, i64 N, i64 nOcc, i64 nPart, double Deg,
                                     std::vector<double>& w, std::vector<double>& wi) {
  w.assign(N, 0.0);
  wi.assign(N, 0.0);
  for (i64 q = 0; q < nOcc && q < N; ++q) {
    if (q < nOcc - nPart) w[q] = 1.0;
    else if (occ_mode == 0) { w[q] = (double)(float)Deg; wi[q] = 8.0; }
    else w[q] = Deg;
  }
}

// SortEnergies_1Spin + the Fermi block (Main.f90:400-412) + GetFock weights.
static void occupations(const forge_ctx* c, const std::vector<double>& band, Occupation& o) {
  const i64 Nk = c->Nk, nb = c->nb, N = nb * Nk;
  std::vector<double> key, val;
  sort_window(c, band, key, val);
  const i64 NP = (c->n / 2 - c->nL) + 1;
  double x = (double)(Nk * NP) + (double)Nk / 2.0 * c->P.nfilling;
  o.nF = (i64)(x >= 0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5));
  o.EF = (key[o.nF - 1] + key[o.nF]) / 2.0;
  o.nOcc = o.nPart = 0;
  for (i64 q = 0; q < N; ++q) {
    if ((key[q] - o.EF) < -c->P.energy_tol) ++o.nOcc;
    else if (std::fabs(key[q] - o.EF) < c->P.energy_tol) { ++o.nOcc; ++o.nPart; }
  }
  o.Deg = 1.0 / (double)o.nPart * (double)(o.nF + o.nPart - o.nOcc);
  std::vector<double> w, wi;
  if (c->opt.occ_mode == 2) {
    w.assign(N, 0.0);
    wi.assign(N, 0.0);
    double kT = c->opt.kT > 0 ? c->opt.kT : 1e-3;
    double lo = key[0] - 50 * kT - 1.0, hi = key[N - 1] + 50 * kT + 1.0, mu = 0;
    for (int it = 0; it < 200; ++it) {
      mu = 0.5 * (lo + hi);
      double s = 0;
      for (i64 q = 0; q < N; ++q) s += 1.0 / (1.0 + std::exp((key[q] - mu) / kT));
      if (s > (double)o.nF) hi = mu; else lo = mu;
    }
    o.EF = mu;
    for (i64 q = 0; q < N; ++q) w[q] = 1.0 / (1.0 + std::exp((key[q] - mu) / kT));
  } else {
    zero_temperature_weights(c->opt.occ_mode, N, o.nOcc, o.nPart, o.Deg, w, wi);
  }
  weighted_states(c, val, w, wi, o);
}

extern "C" int forge_getfock(forge_ctx* c, forge_z* zFock, const forge_z* zSorted,
                             const int64_t* nSortedMomenta, double DegFactor, int64_t OccStates,
                             int64_t PartOccStates) {
  if (!c || !zFock || !zSorted || !nSortedMomenta) { set_error("forge_getfock: null argument"); return -2; }
  const i64 n = c->n, nnc = c->nnc, N = c->nb * c->Nk;
  if (OccStates < 0 || OccStates > N || PartOccStates < 0 || PartOccStates > OccStates) {
    set_error("forge_getfock: invalid state counts");
    return -2;
  }
  const size_t nn = (size_t)n * n;
  std::vector<double> wph;
  for (i64 s = 0; s < OccStates; ++s) {
    std::complex<double> w = 1.0;
    if (s >= OccStates - PartOccStates)
      w = c->opt.occ_mode == 0 ? std::complex<double>((double)(float)DegFactor, 8.0)
                               : std::complex<double>(DegFactor, 0.0);
    i64 i1 = nSortedMomenta[s], i2 = nSortedMomenta[s + N];
    for (i64 cc = 0; cc < nnc; ++cc) {
      double a = -2.0 * PI * (double)(i1 * c->u1[cc] + i2 * c->u1[cc] + i2 * c->u2[cc]) / (double)c->numk;
      std::complex<double> z = std::complex<double>(std::cos(a), std::sin(a)) * w;
      wph.push_back(z.real());
      wph.push_back(z.imag());
    }
  }
  FORGE_HIP(hipMemcpyAsync(c->rho_new, zFock, sizeof(z2) * nnc * nn, hipMemcpyHostToDevice, c->s));
  if (OccStates > 0) {
    FORGE_HIP(hipMemcpyAsync(c->Psi, zSorted, sizeof(z2) * n * OccStates, hipMemcpyHostToDevice, c->s));
    FORGE_HIP(hipMemcpyAsync(c->wph, wph.data(), sizeof(double) * wph.size(), hipMemcpyHostToDevice, c->s));
    CK(window_density(n, (int)nnc, OccStates, c->Psi, c->wph, c->rho_new, 1.0, c->s));
  }
  FORGE_HIP(hipMemcpyAsync(zFock, c->rho_new, sizeof(z2) * nnc * nn, hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipStreamSynchronize(c->s));
  c->scf_ready = false;
  return 0;
}

// ---------------------------------------------------------------- C3 entry points

static int c3_window(forge_ctx* c, i64 S, const std::vector<i64>& cols,
                     const std::vector<i64>& kof, const double* wph, const z2* src,
                     double scale);

extern "C" int64_t forge_c3_pairs(const double* Coords, int64_t ndim, int64_t nlayers,
                                  int64_t* nC3pairs) {
  if (!Coords || !nC3pairs || ndim < 2 || nlayers < 1 || ndim % nlayers != 0 ||
      (ndim / nlayers) % 2 != 0) {
    set_error("forge_c3_pairs: invalid argument");
    return -2;
  }
  const i64 cnt = c3_pairs(Coords, ndim, nlayers, nC3pairs);
  for (i64 i = 0; i < ndim; ++i)
    if (nC3pairs[i] >= 0) nC3pairs[i] += 1;  // Fortran numbering; unmatched sites stay -1
  return cnt;
}

extern "C" int forge_solve_c3(forge_ctx* c, forge_z* zFockBulk, const forge_z* zFock,
                              const double* Potential, double* Bands, forge_z* zEigenvectors,
                              const int64_t* RotateLayers, const int64_t* nC3pairs) {
  if (!c || !zFockBulk || !zFock || !Potential || !Bands || !zEigenvectors) {
    set_error("forge_solve_c3: null argument");
    return -2;
  }
  CK(c3_configure(c, RotateLayers, nC3pairs, "forge_solve_c3"));
  const i64 n = c->n, nnc = c->nnc, Nk = c->Nk, nb = c->nb;
  const size_t nn = (size_t)n * n;
  FORGE_HIP(hipMemcpyAsync(c->rho, zFock, sizeof(z2) * nnc * nn, hipMemcpyHostToDevice, c->s));
  FORGE_HIP(hipMemcpyAsync(c->db, Potential, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
  std::vector<double> hb;
  int f = solve_device_c3(c, c->rho, c->db, hb, false);
  if (f < 0) return f;
  FORGE_HIP(hipMemcpyAsync(zFockBulk, c->rho_new, sizeof(z2) * nnc * nn, hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipMemcpyAsync(zEigenvectors, c->Cwin, sizeof(z2) * n * nb * Nk, hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipStreamSynchronize(c->s));
  std::memcpy(Bands, hb.data(), sizeof(double) * nb * Nk);
  c->scf_ready = false;
  return f;
}

extern "C" int forge_getfock_c3(forge_ctx* c, forge_z* zFock, const forge_z* zSorted,
                                const int64_t* nSortedMomenta, double DegFactor, int64_t OccStates,
                                int64_t PartOccStates, const int64_t* RotateLayers,
                                const int64_t* nC3pairs) {
  if (!c || !zFock || !zSorted || !nSortedMomenta) { set_error("forge_getfock_c3: null argument"); return -2; }
  const i64 n = c->n, nnc = c->nnc, numk = c->numk, N = c->nb * c->Nk;
  if (OccStates < 0 || OccStates > N || PartOccStates < 0 || PartOccStates > OccStates) {
    set_error("forge_getfock_c3: invalid state counts");
    return -2;
  }
  for (i64 s = 0; s < OccStates; ++s)
    if (nSortedMomenta[s] < 0 || nSortedMomenta[s] >= numk || nSortedMomenta[s + N] < 0 ||
        nSortedMomenta[s + N] >= numk) {
      set_error("forge_getfock_c3: momentum component out of range");
      return -2;
    }
  CK(c3_configure(c, RotateLayers, nC3pairs, "forge_getfock_c3"));
  const size_t nn = (size_t)n * n;
  std::vector<double> wph;
  std::vector<i64> cols(OccStates), kof(OccStates);
  for (i64 s = 0; s < OccStates; ++s) {
    std::complex<double> w = 1.0;
    if (s >= OccStates - PartOccStates)
      w = c->opt.occ_mode == 0 ? std::complex<double>((double)(float)DegFactor, 8.0)
                               : std::complex<double>(DegFactor, 0.0);
    i64 i1 = nSortedMomenta[s], i2 = nSortedMomenta[s + N];
    cols[s] = s;
    kof[s] = i1 * numk + i2;
    for (i64 cc = 0; cc < nnc; ++cc) {
      double a = -2.0 * PI * (double)(i1 * c->u1[cc] + i2 * c->u1[cc] + i2 * c->u2[cc]) / (double)numk;
      std::complex<double> z = std::complex<double>(std::cos(a), std::sin(a)) * w;
      wph.push_back(z.real());
      wph.push_back(z.imag());
    }
  }
  FORGE_HIP(hipMemcpyAsync(c->rho_new, zFock, sizeof(z2) * nnc * nn, hipMemcpyHostToDevice, c->s));
  if (OccStates > 0)
    FORGE_HIP(hipMemcpyAsync(c->Psi, zSorted, sizeof(z2) * n * OccStates, hipMemcpyHostToDevice, c->s));
  CK(c3_window(c, OccStates, cols, kof, wph.data(), c->Psi, 1.0));
  FORGE_HIP(hipMemcpyAsync(zFock, c->rho_new, sizeof(z2) * nnc * nn, hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipStreamSynchronize(c->s));
  c->scf_ready = false;
  return 0;
}

// ---------------------------------------------------------------- valley projection

// ValleyGeom from RotateLayers (RotMatrix of Main.f90:133 from the cell order:
// aMoire = ndim / (2 nlayers)) and the workspace, once per context.
static int valley_configure(forge_ctx* c, const int64_t* rot, const char* who) {
  const i64 n = c->n, nl = c->P.nlayers, Nk = c->Nk;
  if (!rot) { set_error(std::string(who) + ": null RotateLayers"); return -2; }
  if (nl < 1 || nl > VALLEY_MAXL || n % (2 * nl) != 0) {
    set_error(std::string(who) + ": layers must split ndim into two equal sublattice halves");
    return -2;
  }
  ValleyGeom g{};
  g.n = n;
  g.per = n / nl;
  g.nl = (int)nl;
  g.nnc = (int)c->nnc;
  const double aM = (double)(n / (2 * nl));
  const double cs = 1.0 - 1.0 / (2.0 * aM);
  const double co = std::cos(0.5 * std::acos(cs)), sn = std::sin(0.5 * std::acos(cs));
  const double Rp[4] = {co, sn, -sn, co}, Rm[4] = {co, -sn, sn, co};
  for (int q = 0; q < 4; ++q) { g.Rp[q] = Rp[q]; g.Rm[q] = Rm[q]; }
  for (i64 l = 0; l < nl; ++l) g.rot[l] = (signed char)((rot[l] == 1 || rot[l] == -1) ? rot[l] : 0);
  g.cells = c->g.cells;
  g.X = c->X;
  g.Y = c->X + n;
  c->vg = g;
  if (!c->valley_ws) CK(dalloc(&c->valley_ws, valley_workspace_bytes(n, c->nb, (int)Nk)));
  c->h_vk.assign(2 * Nk, 0.0);
  for (i64 k = 0; k < Nk; ++k) {
    const i64 i1 = c->nMomComp[k], i2 = c->nMomComp[k + Nk];
    c->h_vk[2 * k] = c->MomVal[i1 + c->numk * i2];
    c->h_vk[2 * k + 1] = c->MomVal[i1 + c->numk * i2 + Nk];
  }
  c->valley_ready = true;
  return 0;
}

// c->Cwin (the window eigenvectors of all Nk momenta) <- ValleyTransform; the
// valley operator is built in the scratch matrix after rho_new's cells.
static int valley_transform_cwin(forge_ctx* c) {
  const size_t nn = (size_t)c->n * c->n;
  return valley_transform(c->vg, c->h_vk.data(), (int)c->Nk, c->nb, c->Cwin,
                          c->rho_new + c->nnc * nn, c->valley_ws, c->s);
}

extern "C" int forge_valley_phase(forge_ctx* c, const double* vk, const int64_t* RotateLayers,
                                  forge_z* zV) {
  if (!c || !vk || !zV) { set_error("forge_valley_phase: null argument"); return -2; }
  CK(valley_configure(c, RotateLayers, "forge_valley_phase"));
  const size_t nn = (size_t)c->n * c->n;
  z2* V = c->rho_new + c->nnc * nn;
  CK(valley_phase(c->vg, vk[0], vk[1], V, c->s));
  FORGE_HIP(hipMemcpyAsync(zV, V, sizeof(z2) * nn, hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipStreamSynchronize(c->s));
  c->scf_ready = false;
  return 0;
}

extern "C" int forge_valley_transform(forge_ctx* c, forge_z* zEigenvectors, const int64_t* RotateLayers) {
  if (!c || !zEigenvectors) { set_error("forge_valley_transform: null argument"); return -2; }
  CK(valley_configure(c, RotateLayers, "forge_valley_transform"));
  const size_t bytes = sizeof(z2) * c->n * c->nb * c->Nk;
  FORGE_HIP(hipMemcpyAsync(c->Cwin, zEigenvectors, bytes, hipMemcpyHostToDevice, c->s));
  CK(valley_transform_cwin(c));
  FORGE_HIP(hipMemcpyAsync(zEigenvectors, c->Cwin, bytes, hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipStreamSynchronize(c->s));
  c->scf_ready = false;
  return 0;
}

extern "C" int forge_scf_set_valley(forge_ctx* c, const int64_t* RotateLayers) {
  if (!c) { set_error("forge_scf_set_valley: null context"); return -2; }
  if (!RotateLayers) {
    c->valley = false;
    return 0;
  }
  CK(valley_configure(c, RotateLayers, "forge_scf_set_valley"));
  c->valley = true;
  return 0;
}

// ---------------------------------------------------------------- T / C2 / C2T

extern "C" int64_t forge_c2_pairs(const double* Coords, int64_t ndim, int64_t nlayers,
                                  int64_t* nC2pairs) {
  if (!Coords || !nC2pairs || ndim < 2 || nlayers < 1 || ndim % nlayers != 0 ||
      (ndim / nlayers) % 2 != 0) {
    set_error("forge_c2_pairs: invalid argument");
    return -2;
  }
  const i64 cnt = c2_pairs(Coords, ndim, nlayers, nC2pairs);
  for (i64 i = 0; i < ndim; ++i)
    if (nC2pairs[i] >= 0) nC2pairs[i] += 1;
  return cnt;
}

// C2 partners (1-based) onto the device: an involution of 1..ndim.
static int c2_configure(forge_ctx* c, int sym, const int64_t* pairs, const char* who) {
  if (sym < 0 || sym > 7) { set_error(std::string(who) + ": flags must combine 1 (T), 2 (C2), 4 (C2T)"); return -2; }
  if (!(sym & 6)) return 0;
  if (!pairs) { set_error(std::string(who) + ": C2 or C2T needs nC2pairs"); return -2; }
  const i64 n = c->n;
  std::vector<int> p(n);
  for (i64 i = 0; i < n; ++i) {
    const int64_t j = pairs[i] - 1;
    if (j < 0 || j >= n) { set_error(std::string(who) + ": nC2pairs out of range"); return -2; }
    p[i] = (int)j;
  }
  for (i64 i = 0; i < n; ++i)
    if (p[p[i]] != i) { set_error(std::string(who) + ": nC2pairs is not an involution"); return -2; }
  if (!c->c2_perm) CK(dalloc((void**)&c->c2_perm, sizeof(int) * n));
  FORGE_HIP(hipMemcpy(c->c2_perm, p.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  return 0;
}

extern "C" int forge_symmetrize(forge_ctx* c, forge_z* zFock, int32_t flags, const int64_t* nC2pairs) {
  if (!c || !zFock) { set_error("forge_symmetrize: null argument"); return -2; }
  CK(c2_configure(c, flags, nC2pairs, "forge_symmetrize"));
  const i64 n = c->n, nnc = c->nnc;
  const size_t nn = (size_t)n * n;
  FORGE_HIP(hipMemcpyAsync(c->rho_new, zFock, sizeof(z2) * nnc * nn, hipMemcpyHostToDevice, c->s));
  CK(symmetrize(n, (int)nnc, c->rho_new, c->rho_new + nnc * nn, flags, c->c2_perm, c->s));
  FORGE_HIP(hipMemcpyAsync(zFock, c->rho_new, sizeof(z2) * nnc * nn, hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipStreamSynchronize(c->s));
  c->scf_ready = false;
  return 0;
}

extern "C" int forge_scf_set_symmetry(forge_ctx* c, int32_t flags, const int64_t* nC2pairs) {
  if (!c) { set_error("forge_scf_set_symmetry: null context"); return -2; }
  CK(c2_configure(c, flags, nC2pairs, "forge_scf_set_symmetry"));
  c->sym = flags;
  return 0;
}

extern "C" int forge_scf_set_c3(forge_ctx* c, const int64_t* RotateLayers, const int64_t* nC3pairs) {
  if (!c) { set_error("forge_scf_set_c3: null context"); return -2; }
  if (!RotateLayers && !nC3pairs) {
    c->c3 = false;
    return 0;
  }
  CK(c3_configure(c, RotateLayers, nC3pairs, "forge_scf_set_c3"));
  c->c3 = true;
  return 0;
}

// Energy pieces of a = rho (dens da) and optionally b = rho_new (dens db):
// pieces[0..4] = K_a, K_b, F_aa, F_ab, F_bb; [5..7] = H_aa, H_ab, H_bb
// (d.L.d), [8..10] = U_aa, U_ab, U_bb; [11..12] bias terms of a, b.
static int energy_pieces(forge_ctx* c, const z2* a, const z2* b, double* pieces) {
  const i64 n = c->n;
  const size_t nn = (size_t)n * n;
  CK(ensure_longrange(c));
  const int nblk = 1024;
  CK(energy_pass(nn, (int)c->nnc, a, b, c->V, c->T, c->part, nblk, c->e5, c->s));
  CK(hartree_matvec(n, c->L, c->da, b ? c->db : nullptr, c->ya, b ? c->yb : nullptr, c->s));
  hipLaunchKernelGGL(k_dots, dim3(1), dim3(256), 0, c->s, n, c->da, b ? c->db : nullptr, c->ya,
                     b ? c->yb : nullptr, c->dots);
  FORGE_LAUNCH();
  hipLaunchKernelGGL(k_bias_dot, dim3(1), dim3(256), 0, c->s, n, c->da, b ? c->db : nullptr, c->bias,
                     c->dots + 8);
  FORGE_LAUNCH();
  double* h5 = c->hpin + n * c->Kb + 2 * n;  // 5 energy sums, then 10 dot products
  double* hd = h5 + 5;
  FORGE_HIP(hipMemcpyAsync(h5, c->e5, sizeof(double) * 5, hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipMemcpyAsync(hd, c->dots, sizeof(double) * 10, hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipStreamSynchronize(c->s));
  for (int q = 0; q < 5; ++q) pieces[q] = h5[q];
  pieces[5] = hd[0];
  pieces[6] = hd[1];
  pieces[7] = hd[2];
  pieces[8] = hd[3];
  pieces[9] = hd[4];
  pieces[10] = hd[5];
  pieces[11] = hd[8];
  pieces[12] = hd[9];
  return 0;
}

// energies (Ekin, EF, EH, EU) of (1-l) a + l b from the pieces
static void energies_at(const double* pc, double l, double* e) { scf_energies_at(pc, l, e); }

extern "C" int forge_energies(forge_ctx* c, const forge_z* zFock, const double* dens, double* out4) {
  if (!c || !zFock || !dens || !out4) { set_error("forge_energies: null argument"); return -2; }
  const size_t nn = (size_t)c->n * c->n;
  FORGE_HIP(hipMemcpyAsync(c->rho_new, zFock, sizeof(z2) * c->nnc * nn, hipMemcpyHostToDevice, c->s));
  FORGE_HIP(hipMemcpyAsync(c->da, dens, sizeof(double) * c->n, hipMemcpyHostToDevice, c->s));
  double pc[13];
  CK(energy_pieces(c, c->rho_new, nullptr, pc));
  energies_at(pc, 0.0, out4);
  c->scf_ready = false;
  return 0;
}

extern "C" int forge_hartree_hubbard(forge_ctx* c, const double* dens1, const double* dens2,
                                     double* out2) {
  if (!c || !dens1 || !dens2 || !out2) { set_error("forge_hartree_hubbard: null argument"); return -2; }
  const i64 n = c->n;
  CK(ensure_longrange(c));
  FORGE_HIP(hipMemcpyAsync(c->da, dens1, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
  FORGE_HIP(hipMemcpyAsync(c->db, dens2, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
  CK(hartree_matvec(n, c->L, c->da, c->db, c->ya, c->yb, c->s));
  hipLaunchKernelGGL(k_dots, dim3(1), dim3(256), 0, c->s, n, c->da, c->db, c->ya, c->yb, c->dots);
  FORGE_LAUNCH();
  double hd[6];
  FORGE_HIP(hipMemcpyAsync(hd, c->dots, sizeof(hd), hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipStreamSynchronize(c->s));
  // (d1 + d2).L.(d1 + d2) / 2 with L symmetric; d1.d2
  out2[0] = 0.5 * (hd[0] + 2.0 * hd[1] + hd[2]);
  out2[1] = hd[4];
  c->scf_ready = false;
  return 0;
}

extern "C" int forge_scf_set_options(forge_ctx* c, const forge_scf_options* opt) {
  if (!c || !opt) { set_error("forge_scf_set_options: null argument"); return -2; }
  if (opt->occ_mode < 0 || opt->occ_mode > 2 || opt->mix_mode < 0 || opt->mix_mode > 1 ||
      (opt->mix_mode == 1 && !(opt->beta > 0.0 && opt->beta <= 1.0)) ||
      (opt->occ_mode == 2 && !(opt->kT > 0.0))) {
    set_error("forge_scf_set_options: invalid option values");
    return -2;
  }
  if (opt->precision < 0 || opt->precision > 1 || !(opt->switch_de >= 0.0)) {
    set_error("forge_scf_set_options: precision must be 0 or 1 and switch_de >= 0");
    return -2;
  }
  if (opt->graph < 0 || opt->graph > 3 || (opt->graph >= 2 && opt->timing)) {
    set_error("forge_scf_set_options: graph must be 0..3, and 2 or 3 without timing");
    return -2;
  }
  if (opt->graph >= 2 && !c->dband) {
    CK(dalloc((void**)&c->dband, sizeof(double) * c->nb * c->Nk));
    CK(dalloc((void**)&c->dS, sizeof(i64)));
    CK(dalloc((void**)&c->docc, sizeof(double) * 4));
    CK(dalloc((void**)&c->dst, sizeof(double) * 8));
    CK(dalloc((void**)&c->drow, sizeof(double) * 12));
  }
  // the iteration graphs hold the options as kernel arguments
  for (auto& g : c->it_graph)
    if (g) {
      FORGE_HIP(hipGraphExecDestroy(g));
      g = nullptr;
    }
  if (opt->reduction < 0 || opt->reduction > 2) {
    set_error("forge_scf_set_options: reduction must be 0 (by order), 1 (one-stage) or 2 (two-stage)");
    return -2;
  }
  if (reduction_mode(opt->reduction, c->n) == 2 && !c->ts_ws) {
    const size_t bytes = twostage_workspace_bytes(c->n, c->Kb);
    size_t fr = 0, tot = 0;
    FORGE_HIP(hipMemGetInfo(&fr, &tot));
    if (bytes + ((size_t)256 << 20) > fr) {
      set_error("forge_scf_set_options: not enough device memory for the two-stage reduction");
      return -4;
    }
    CK(dalloc(&c->ts_ws, bytes));
    CK(twostage_bind(c->ts, c->ts_ws, c->n, c->Kb));
  }
  if (opt->precision == 1 && !c->S) {
    const size_t bytes = sizeof(float2) * (size_t)c->ldH * c->n * c->Kb;
    size_t fr = 0, tot = 0;
    FORGE_HIP(hipMemGetInfo(&fr, &tot));
    if (bytes + ((size_t)256 << 20) > fr) {
      set_error("forge_scf_set_options: not enough device memory for the FP32 copy of precision 1");
      return -4;
    }
    CK(dalloc((void**)&c->S, bytes));
  }
  c->opt = *opt;
  return 0;
}

// GetFock_C3's window columns (HartreeFock.f90:564-881): a state at a
// C3-invariant momentum becomes three columns, itself and its two images,
// each with a third of its weight; every other state is copied as is.
// cols[s] is the source column of state s, wph its nnc weights.
static int c3_window(forge_ctx* c, i64 S, const std::vector<i64>& cols,
                     const std::vector<i64>& kof, const double* wph, const z2* src,
                     double scale) {
  const i64 n = c->n, nnc = c->nnc;
  c->c3_psi.clear();
  c->c3_wph_h.clear();
  for (i64 s = 0; s < S; ++s) {
    const i64 k = kof[s];
    const int copies = c->c3_inv[k] ? 3 : 1;
    for (int r = 0; r < copies; ++r) {
      const int set = r == 0 ? -1 : (int)(6 * k + 1 + r);  // conj(f), conj(u)
      c->c3_psi.push_back(C3Col{cols[s], (i64)c->c3_psi.size(), r, set});
      for (i64 q = 0; q < 2 * nnc; ++q) c->c3_wph_h.push_back(wph[s * nnc * 2 + q] / (double)copies);
    }
  }
  const i64 S3 = (i64)c->c3_psi.size();
  if (S3 > c->c3_cap) { set_error("c3_window: more window columns than allocated"); return -2; }
  if (S3 > 0) {
    FORGE_HIP(hipMemcpyAsync(c->c3_dpsi, c->c3_psi.data(), sizeof(C3Col) * S3, hipMemcpyHostToDevice, c->s));
    FORGE_HIP(hipMemcpyAsync(c->c3_wph, c->c3_wph_h.data(), sizeof(double) * c->c3_wph_h.size(),
                             hipMemcpyHostToDevice, c->s));
    CK(c3_columns(c->c3g, S3, c->c3_dpsi, src, c->c3_Psi, c->s));
  }
  CK(window_density(n, (int)nnc, S3, c->c3_Psi, c->c3_wph, c->rho_new, scale, c->s));
  // the descriptor vectors are reused by the next call: wait for the copies
  FORGE_HIP(hipStreamSynchronize(c->s));
  return 0;
}

static int valley_transform_cwin(forge_ctx* c);

// window density into c->rho_new (which holds the bulk): rho_new = (bulk + win) / Nk.
// in_loop: an SCF iteration (the valley projection applies only there; the
// initial states of Main.f90:238-288 and of the guesses use the plain states)
static int window_of(forge_ctx* c, const Occupation& occ, bool in_loop);

static int window_step(forge_ctx* c, const std::vector<double>& band, Occupation& occ,
                       bool in_loop) {
  occupations(c, band, occ);
  return window_of(c, occ, in_loop);
}

// the window density of the weighted states occ (see window_step)
static int window_of(forge_ctx* c, const Occupation& occ, bool in_loop) {
  if (c->valley && in_loop) {
    // Main.f90:469-494: GetFock of the states, ValleyTransform, GetFock of
    // the transformed states (same energies, so the same sorted columns),
    // rho_new = (bulk + (win + win_valley) / 2) / Nk; with the C3-reduced
    // k-set both window densities are GetFock_C3's (Main.f90:473-476,
    // 487-490), on the window states of every momentum (the partners' are
    // the permuted copies Solve_C3 wrote)
    const i64 n = c->n, S = (i64)occ.idx.size();
    std::vector<double> half(occ.wph);
    for (double& x : half) x *= 0.5;
    if (c->c3) {
      std::vector<i64> cols(S), kof(S);
      for (i64 s = 0; s < S; ++s) {
        cols[s] = occ.idx[s];
        kof[s] = occ.idx[s] / c->nb;
      }
      CK(c3_window(c, S, cols, kof, half.data(), c->Cwin, 1.0));
      CK(valley_transform_cwin(c));
      return c3_window(c, S, cols, kof, half.data(), c->Cwin, 1.0 / (double)c->Nk);
    }
    if (S > 0) {
      FORGE_HIP(hipMemcpyAsync(c->gidx, occ.idx.data(), sizeof(int) * S, hipMemcpyHostToDevice, c->s));
      FORGE_HIP(hipMemcpyAsync(c->wph, half.data(), sizeof(double) * half.size(), hipMemcpyHostToDevice, c->s));
      hipLaunchKernelGGL(k_gather_cols, dim3(cdiv(n * S, 256)), dim3(256), 0, c->s, n, S, c->gidx,
                         c->Cwin, c->Psi);
      FORGE_LAUNCH();
    }
    CK(window_density(n, (int)c->nnc, S, c->Psi, c->wph, c->rho_new, 1.0, c->s));
    CK(valley_transform_cwin(c));
    if (S > 0) {
      hipLaunchKernelGGL(k_gather_cols, dim3(cdiv(n * S, 256)), dim3(256), 0, c->s, n, S, c->gidx,
                         c->Cwin, c->Psi);
      FORGE_LAUNCH();
    }
    CK(window_density(n, (int)c->nnc, S, c->Psi, c->wph, c->rho_new, 1.0 / (double)c->Nk, c->s));
    // half lives on the host until the copies above have run
    FORGE_HIP(hipStreamSynchronize(c->s));
    return 0;
  }
  if (c->c3) {
    const i64 S = (i64)occ.idx.size();
    std::vector<i64> cols(S), kof(S);
    for (i64 s = 0; s < S; ++s) {
      cols[s] = occ.idx[s];
      kof[s] = occ.idx[s] / c->nb;
    }
    return c3_window(c, S, cols, kof, occ.wph.data(), c->Cwin, 1.0 / (double)c->Nk);
  }
  const i64 n = c->n, S = (i64)occ.idx.size();
  if (S > 0) {
    FORGE_HIP(hipMemcpyAsync(c->gidx, occ.idx.data(), sizeof(int) * S, hipMemcpyHostToDevice, c->s));
    FORGE_HIP(hipMemcpyAsync(c->wph, occ.wph.data(), sizeof(double) * occ.wph.size(), hipMemcpyHostToDevice, c->s));
    hipLaunchKernelGGL(k_gather_cols, dim3(cdiv(n * S, 256)), dim3(256), 0, c->s, n, S, c->gidx,
                       c->Cwin, c->Psi);
    FORGE_LAUNCH();
  }
  CK(window_density(n, (int)c->nnc, S, c->Psi, c->wph, c->rho_new, 1.0 / (double)c->Nk, c->s));
  return 0;
}

static int init_state_of_rho(forge_ctx* c);

extern "C" int forge_scf_init(forge_ctx* c, const forge_z* rho_start) {
  if (!c) { set_error("forge_scf_init: null context"); return -2; }
  const i64 n = c->n, nnc = c->nnc;
  const size_t nn = (size_t)n * n;
  std::fill(c->tm, c->tm + FORGE_NTIMINGS, 0.0);
  c->mixed_done = false;
  int fails = 0;
  if (rho_start) {
    FORGE_HIP(hipMemcpyAsync(c->rho, rho_start, sizeof(z2) * nnc * nn, hipMemcpyHostToDevice, c->s));
  } else {
    // first diagonalisation: rho = 0, Potential = 0 (Main.f90:238-282)
    FORGE_HIP(hipMemsetAsync(c->pot, 0, sizeof(double) * n, c->s));
    std::vector<double> band;
    fails = c->c3 ? solve_device_c3(c, nullptr, c->pot, band, false)
                  : solve_device(c, nullptr, c->pot, band, false);
    if (fails < 0) return fails;
    Occupation occ;
    CK(window_step(c, band, occ, false));
    FORGE_HIP(hipMemcpyAsync(c->rho, c->rho_new, sizeof(z2) * nnc * nn, hipMemcpyDeviceToDevice, c->s));
  }
  CK(init_state_of_rho(c));
  return fails;
}

// energies of c->rho and the potential it defines (Main.f90:290-346)
static int init_state_of_rho(forge_ctx* c) {
  const i64 n = c->n;
  CK(site_density(n, c->rho, c->da, c->s));
  double pc[13];
  CK(energy_pieces(c, c->rho, nullptr, pc));
  energies_at(pc, 0.0, c->e);
  c->Etot = total_energy(c->P, c->e);
  hipLaunchKernelGGL(k_potential, dim3(cdiv(n, 256)), dim3(256), 0, c->s, n, c->ya, nullptr, c->da,
                     nullptr, 0.0, c->P.alphaH, c->P.U, c->pot, nullptr);
  FORGE_LAUNCH();
  FORGE_HIP(hipStreamSynchronize(c->s));
  c->it = 0;
  c->scf_ready = true;
  c->dev_state = false;
  // every solve starts each phase on the host path, so that runs with and
  // without the iteration graph take the same arithmetic path
  c->dev_warm[0] = c->dev_warm[1] = false;
  return 0;
}

// rho(i, i, 0) = diag_i; rho(row_q, col_q, cell_q) += delta_q (distinct entries)
__global__ void k_apply_guess(i64 n, const double* diag, i64 nupd, const i64* row, const i64* col,
                              const i64* cell, const z2* delta, z2* rho) {
  const i64 t = (i64)blockIdx.x * 256 + threadIdx.x;
  if (t < n) rho[t + n * t] = mk(diag[t], 0.0);
  if (t < nupd) {
    z2* e = rho + (size_t)cell[t] * n * n + row[t] + (size_t)n * col[t];
    *e = *e + delta[t];
  }
}

// one diagonalisation of H[c->rho] at every k; the state becomes its
// occupied density
static int guess_diagonalisation(forge_ctx* c) {
  const size_t tot = (size_t)c->nnc * c->n * c->n;
  CK(init_state_of_rho(c));
  std::vector<double> band;
  const int fails = c->c3 ? solve_device_c3(c, c->rho, c->pot, band, false)
                          : solve_device(c, c->rho, c->pot, band, false);
  if (fails < 0) return fails;
  Occupation occ;
  CK(window_step(c, band, occ, false));
  FORGE_HIP(hipMemcpyAsync(c->rho, c->rho_new, sizeof(z2) * tot, hipMemcpyDeviceToDevice, c->s));
  CK(init_state_of_rho(c));
  return fails;
}

// Applies the learned guess to the current state on the device: diagonal
// values and distinct (row, col, cell) increments.
static int apply_guess(forge_ctx* c, const char* who, const double* diag, int64_t nupd,
                       const int64_t* row, const int64_t* col, const int64_t* cell,
                       const forge_z* delta) {
  if (!c || !diag || nupd < 0 || (nupd > 0 && (!row || !col || !cell || !delta))) {
    set_error(std::string(who) + ": invalid argument");
    return -2;
  }
  if (!c->scf_ready) { set_error(std::string(who) + ": call forge_scf_init first"); return -5; }
  const i64 n = c->n;
  for (i64 q = 0; q < nupd; ++q)
    if (row[q] < 0 || row[q] >= n || col[q] < 0 || col[q] >= n || cell[q] < 0 || cell[q] >= c->nnc) {
      set_error(std::string(who) + ": entry out of range");
      return -2;
    }
  std::fill(c->tm, c->tm + FORGE_NTIMINGS, 0.0);
  c->mixed_done = false;
  double* dd = nullptr;
  i64* ix = nullptr;
  z2* dv = nullptr;
  const i64 m = std::max<i64>(nupd, 1);
  int rc = [&]() -> int {
    CK(dalloc((void**)&dd, sizeof(double) * n));
    CK(dalloc((void**)&ix, sizeof(i64) * 3 * m));
    CK(dalloc((void**)&dv, sizeof(z2) * m));
    FORGE_HIP(hipMemcpyAsync(dd, diag, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
    if (nupd > 0) {
      FORGE_HIP(hipMemcpyAsync(ix, row, sizeof(i64) * nupd, hipMemcpyHostToDevice, c->s));
      FORGE_HIP(hipMemcpyAsync(ix + m, col, sizeof(i64) * nupd, hipMemcpyHostToDevice, c->s));
      FORGE_HIP(hipMemcpyAsync(ix + 2 * m, cell, sizeof(i64) * nupd, hipMemcpyHostToDevice, c->s));
      FORGE_HIP(hipMemcpyAsync(dv, delta, sizeof(z2) * nupd, hipMemcpyHostToDevice, c->s));
    }
    hipLaunchKernelGGL(k_apply_guess, dim3((unsigned)cdiv(std::max<i64>(n, nupd), 256)), dim3(256), 0,
                       c->s, n, dd, nupd, ix, ix + m, ix + 2 * m, dv, c->rho);
    FORGE_LAUNCH();
    FORGE_HIP(hipStreamSynchronize(c->s));
    return 0;
  }();
  if (dd) (void)hipFree(dd);
  if (ix) (void)hipFree(ix);
  if (dv) (void)hipFree(dv);
  return rc;
}

extern "C" int forge_scf_init_guess_update(forge_ctx* c, const double* diag, int64_t nupd,
                                           const int64_t* row, const int64_t* col,
                                           const int64_t* cell, const forge_z* delta) {
  CK(apply_guess(c, "forge_scf_init_guess_update", diag, nupd, row, col, cell, delta));
  return guess_diagonalisation(c);
}

// The guess itself as the start state (its potential and energies; no
// diagonalisation): the first SCF iteration then solves H[guess] and mixes.
extern "C" int forge_scf_init_update(forge_ctx* c, const double* diag, int64_t nupd,
                                     const int64_t* row, const int64_t* col, const int64_t* cell,
                                     const forge_z* delta) {
  CK(apply_guess(c, "forge_scf_init_update", diag, nupd, row, col, cell, delta));
  return init_state_of_rho(c);
}

extern "C" int forge_scf_init_guess(forge_ctx* c, const forge_z* rho_guess) {
  if (!c || !rho_guess) { set_error("forge_scf_init_guess: null argument"); return -2; }
  const i64 n = c->n, nnc = c->nnc;
  const size_t nn = (size_t)n * n;
  std::fill(c->tm, c->tm + FORGE_NTIMINGS, 0.0);
  c->mixed_done = false;
  // H[rho_guess]: its potential, then one diagonalisation at every k
  FORGE_HIP(hipMemcpyAsync(c->rho, rho_guess, sizeof(z2) * nnc * nn, hipMemcpyHostToDevice, c->s));
  return guess_diagonalisation(c);
}

// One SCF iteration of Main.f90:380-600: solve at every k, Fermi level and
// window density, energies, OptimalStep or linear mixing, new potential.
// traj_row (12 values) may be null; *stop: converged (FP64 iteration with
// |dE| < etol).  eig_ready: see solve_device.  Returns D&C failures or < 0.
static int scf_step(forge_ctx* c, double* traj_row, bool timed, bool eig_ready, bool* stop) {
  const i64 n = c->n, nnc = c->nnc;
  const size_t nn = (size_t)n * n;
  *stop = false;
  double* hda = c->hpin + n * c->Kb;  // the site densities of rho and rho_new
  double* hdb = hda + n;
  std::vector<double> band;
  const bool mixed = c->opt.precision == 1 && !c->mixed_done;
  if (mixed) c->tm[12] += 1;
  if (c->c3 && eig_ready) { set_error("scf_step: the C3-reduced k-set is not batched in sweeps"); return -2; }
  int fails = c->c3 ? solve_device_c3(c, c->rho, c->pot, band, timed)
                    : solve_device(c, c->rho, c->pot, band, timed, eig_ready);
  if (fails < 0) return fails;
  if (timed) FORGE_HIP(hipEventRecord(c->ev[10], c->s));
  Occupation occ;
  CK(window_step(c, band, occ, true));
  // symmetry enforcement on the new density (Main.f90:514-534)
  if (c->sym) CK(symmetrize(n, (int)nnc, c->rho_new, c->rho_new + nnc * nn, c->sym, c->c2_perm, c->s));
  if (timed) {
    FORGE_HIP(hipEventRecord(c->ev[11], c->s));
    FORGE_HIP(hipEventSynchronize(c->ev[11]));
    float a;
    FORGE_HIP(hipEventElapsedTime(&a, c->ev[10], c->ev[11]));
    c->tm[4] += a;
    FORGE_HIP(hipEventRecord(c->ev[12], c->s));
  }
  CK(site_density(n, c->rho_new, c->db, c->s));
  double pc[13];
  CK(energy_pieces(c, c->rho, c->rho_new, pc));
  double ein[4] = {c->e[0], c->e[1], c->e[2], c->e[3]};
  const double EtotIn = c->Etot;
  double step;
  double e1[4], eh[4];
  energies_at(pc, 1.0, e1);
  energies_at(pc, 0.5, eh);
  if (c->opt.mix_mode == 0) {
    // OptimalStep, HartreeFock.f90:1162-1208
    step = scf_oda_step(EtotIn, total_energy(c->P, e1), total_energy(c->P, eh), c->P.step_alt);
  } else {
    step = c->opt.beta;
  }
  energies_at(pc, step, c->e);
  c->Etot = total_energy(c->P, c->e);
  // DensityConvergence = max |D_mixed - D_in| (Main.f90:564-571)
  FORGE_HIP(hipMemcpyAsync(hda, c->da, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipMemcpyAsync(hdb, c->db, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
  CK(mix(nnc * nn, c->rho, c->rho_new, step, c->s));
  hipLaunchKernelGGL(k_potential, dim3(cdiv(n, 256)), dim3(256), 0, c->s, n, c->ya, c->yb, c->da,
                     c->db, step, c->P.alphaH, c->P.U, c->pot, c->da);
  FORGE_LAUNCH();
  FORGE_HIP(hipStreamSynchronize(c->s));
  double dconv = 0;
  for (i64 q = 0; q < n; ++q) {
    double dmix = (1.0 - step) * hda[q] + step * hdb[q];
    dconv = std::max(dconv, std::fabs(dmix - hda[q]));
  }
  if (timed) {
    FORGE_HIP(hipEventRecord(c->ev[13], c->s));
    FORGE_HIP(hipEventSynchronize(c->ev[13]));
    float a;
    FORGE_HIP(hipEventElapsedTime(&a, c->ev[12], c->ev[13]));
    c->tm[5] += a;
  }
  ++c->it;
  if (traj_row) {
    double* r = traj_row;
    r[0] = (double)c->it; r[1] = EtotIn; r[2] = 2 * ein[0]; r[3] = 2 * ein[1]; r[4] = ein[2];
    r[5] = ein[3]; r[6] = dconv; r[7] = step; r[8] = EtotIn - c->Etot; r[9] = occ.EF;
    r[10] = (double)occ.nOcc; r[11] = (double)occ.nPart;
  }
  const double dE = std::fabs(EtotIn - c->Etot);
  if (mixed) {
    // precision 1: FP64 from here on once the energy change reaches the
    // switch threshold (default 1e-4 eV); convergence is only declared on
    // FP64 iterations
    const double sw = c->opt.switch_de > 0 ? c->opt.switch_de : 1e-4;
    if (dE < sw || (c->opt.etol > 0 && dE < c->opt.etol)) c->mixed_done = true;
    return fails;
  }
  if (c->opt.etol > 0 && dE < c->opt.etol) *stop = true;
  return fails;
}

// ---------------------------------------------------------------- device-resident iteration
// options.graph 2 / 3: the iteration of scf_step with the sort, Fermi level,
// occupations, OptimalStep and convergence measure on the device
// (scf_device.hip), so that it is one launch sequence; with 3 it is captured
// into a HIP graph per phase (mixed, FP64) on its first use and replayed.
// The first iteration of each phase runs on the host path: it tunes the
// back-transformation's block width and leaves the divide-and-conquer's
// merge descriptors on the device, which the deferred solve then reuses.
// Applies to the plain loop (no C3 set, valley projection or symmetry
// enforcement) with all k-points in one eigensolver batch.
static bool dev_eligible(const forge_ctx* c) {
  return !c->c3 && !c->valley && c->sym == 0 && c->Kb == c->Nk && c->nnc == 4 &&
         c->nb * c->Nk <= OCC_MAX;
}

static int enqueue_device_iteration(forge_ctx* c) {
  const i64 n = c->n, nnc = c->nnc, Nk = c->Nk;
  const size_t nn = (size_t)n * n;
  std::vector<double> unused;
  c->dev_iter = true;
  const int f = solve_device(c, c->rho, c->pot, unused, false);
  c->dev_iter = false;
  if (f < 0) return f;
  OccArgs oa{};
  oa.N = (int)(c->nb * Nk);
  oa.Nk = (int)Nk;
  oa.nb = (int)c->nb;
  oa.nnc = (int)nnc;
  oa.numk = (int)c->numk;
  oa.occ_mode = c->opt.occ_mode;
  {
    const double x = (double)(Nk * ((n / 2 - c->nL) + 1)) + (double)Nk / 2.0 * c->P.nfilling;
    oa.nF = (i64)(x >= 0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5));
  }
  oa.energy_tol = c->P.energy_tol;
  oa.kT = c->opt.kT > 0 ? c->opt.kT : 1e-3;
  oa.band = c->dband;
  oa.ph_cell = c->ph_cell;
  oa.idx = c->gidx;
  oa.wph = c->wph;
  oa.S = c->dS;
  oa.occ = c->docc;
  CK(device_occupations(oa, c->s));
  CK(gather_cols_dev(n, c->nb * Nk, c->dS, c->gidx, c->Cwin, c->Psi, c->s));
  CK(window_density(n, (int)nnc, 0, c->Psi, c->wph, c->rho_new, 1.0 / (double)Nk, c->s, c->dS));
  CK(site_density(n, c->rho_new, c->db, c->s));
  CK(energy_pass(nn, (int)nnc, c->rho, c->rho_new, c->V, c->T, c->part, 1024, c->e5, c->s));
  CK(hartree_matvec(n, c->L, c->da, c->db, c->ya, c->yb, c->s));
  hipLaunchKernelGGL(k_dots, dim3(1), dim3(256), 0, c->s, n, c->da, c->db, c->ya, c->yb, c->dots);
  FORGE_LAUNCH();
  hipLaunchKernelGGL(k_bias_dot, dim3(1), dim3(256), 0, c->s, n, c->da, c->db, c->bias, c->dots + 8);
  FORGE_LAUNCH();
  StepArgs sa{};
  sa.mix_mode = c->opt.mix_mode;
  sa.beta = c->opt.beta;
  sa.step_alt = c->P.step_alt;
  sa.alpha = c->P.alpha;
  sa.alphaH = c->P.alphaH;
  sa.U = c->P.U;
  sa.e5 = c->e5;
  sa.dots = c->dots;
  sa.occ = c->docc;
  sa.st = c->dst;
  sa.row = c->drow;
  CK(scf_step_update(sa, c->s));
  CK(density_convergence(n, c->da, c->db, c->dst, c->drow, c->s));
  CK(mix_dev(nnc * nn, c->rho, c->rho_new, c->dst, c->s));
  CK(potential_dev(n, c->ya, c->yb, c->da, c->db, c->dst, c->P.alphaH, c->P.U, c->pot, c->s));
  return 0;
}

static int scf_step_dev(forge_ctx* c, double* traj_row, bool* stop) {
  const bool mixed = c->opt.precision == 1 && !c->mixed_done;
  const int ph = mixed ? 0 : 1;
  if (!c->dev_warm[ph]) {
    const int f = scf_step(c, traj_row, false, false, stop);
    if (f >= 0) c->dev_warm[ph] = true;
    c->dev_state = false;
    return f;
  }
  *stop = false;
  if (mixed) c->tm[12] += 1;
  c->trd.S = mixed ? c->S : nullptr;
  c->trd.sS = (size_t)c->ldH * c->n;
  c->trd.K = c->Kb;
  if (!c->dev_state) {
    double st[8] = {c->e[0], c->e[1], c->e[2], c->e[3], c->Etot, 0, 0, 0};
    FORGE_HIP(hipMemcpyAsync(c->dst, st, sizeof(st), hipMemcpyHostToDevice, c->s));
    FORGE_HIP(hipStreamSynchronize(c->s));
    c->dev_state = true;
  }
  if (c->opt.graph == 3) {
    if (!c->it_graph[ph]) {
      hipGraph_t g = nullptr;
      FORGE_HIP(hipStreamBeginCapture(c->s, hipStreamCaptureModeRelaxed));
      const int rc = enqueue_device_iteration(c);
      const hipError_t ce = hipStreamEndCapture(c->s, &g);
      if (rc != 0) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (ce != hipSuccess) {
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        return -1;
      }
      const hipError_t ie = hipGraphInstantiate(&c->it_graph[ph], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) {
        c->it_graph[ph] = nullptr;
        set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
        return -1;
      }
    }
    FORGE_HIP(hipGraphLaunch(c->it_graph[ph], c->s));
  } else {
    CK(enqueue_device_iteration(c));
  }
  double row[12], st[8];
  int fails = 0;
  FORGE_HIP(hipMemcpyAsync(row, c->drow, sizeof(row), hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipMemcpyAsync(st, c->dst, sizeof(st), hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipMemcpyAsync(&fails, dc_fail_ptr(c->dc_ws, c->n, c->Kb), sizeof(int),
                           hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipStreamSynchronize(c->s));
  for (int q = 0; q < 4; ++q) c->e[q] = st[q];
  c->Etot = st[4];
  ++c->it;
  row[0] = (double)c->it;
  if (traj_row)
    for (int q = 0; q < 12; ++q) traj_row[q] = row[q];
  const double dE = std::fabs(row[8]);
  if (mixed) {
    const double sw = c->opt.switch_de > 0 ? c->opt.switch_de : 1e-4;
    if (dE < sw || (c->opt.etol > 0 && dE < c->opt.etol)) c->mixed_done = true;
    return fails;
  }
  if (c->opt.etol > 0 && dE < c->opt.etol) *stop = true;
  return fails;
}

extern "C" int forge_scf_run(forge_ctx* c, int64_t max_iter, double* traj, int64_t* iters_done) {
  if (!c || max_iter < 0) { set_error("forge_scf_run: invalid argument"); return -2; }
  if (!c->scf_ready) { set_error("forge_scf_run: call forge_scf_init first"); return -5; }
  std::fill(c->tm, c->tm + FORGE_NTIMINGS, 0.0);
  const bool timed = c->opt.timing != 0;
  int fails = 0;
  i64 done = 0;
  FORGE_HIP(hipEventRecord(c->ev[8], c->s));
  const bool dev = c->opt.graph >= 2;
  if (dev && !dev_eligible(c)) {
    set_error("forge_scf_run: options.graph 2 and 3 apply to the plain loop (no C3 set, valley "
              "projection or symmetry enforcement) with every k-point in one eigensolver batch");
    return -2;
  }
  for (i64 itn = 0; itn < max_iter; ++itn) {
    bool stop = false;
    int f = dev ? scf_step_dev(c, traj ? traj + 12 * itn : nullptr, &stop)
                : scf_step(c, traj ? traj + 12 * itn : nullptr, timed, false, &stop);
    if (f < 0) return f;
    fails += f;
    ++done;
    if (stop) break;
  }
  FORGE_HIP(hipEventRecord(c->ev[9], c->s));
  FORGE_HIP(hipEventSynchronize(c->ev[9]));
  float tt;
  FORGE_HIP(hipEventElapsedTime(&tt, c->ev[8], c->ev[9]));
  c->tm[6] = tt;
  if (iters_done) *iters_done = done;
  return fails;
}

extern "C" int forge_scf_get_density(forge_ctx* c, forge_z* rho) {
  if (!c || !rho) { set_error("forge_scf_get_density: null argument"); return -2; }
  FORGE_HIP(hipMemcpyAsync(rho, c->rho, sizeof(z2) * c->nnc * c->n * c->n, hipMemcpyDeviceToHost, c->s));
  FORGE_HIP(hipStreamSynchronize(c->s));
  return 0;
}

extern "C" int forge_scf_timings(forge_ctx* c, double* out24) {
  if (!c || !out24) { set_error("forge_scf_timings: null argument"); return -2; }
  for (int q = 0; q < FORGE_NTIMINGS; ++q) out24[q] = c->tm[q];
  return 0;
}

// ---------------------------------------------------------------- two spins (numS = 2)
// The SCF loop of Main.f90:236-679 with numS = 2 on the device: two contexts
// of the same system, one per spin, each holding its spin's density, bulk
// and window states, potential and eigensolver batch.  Per iteration:
// Solve (or Solve_C3) for each spin with its own potential, SortEnergies_1Spin
// per spin, SortEnergies_2Spins and the two-spin Fermi block
// (Main.f90:416-458), GetFock per spin with its own counts and the shared
// DegFactor, OptimalStep on the two-spin energy (HartreeFock.f90:1148-1227:
// kinetic and Fock energies summed over the spins, the Hartree energy of the
// total density and the Hubbard energy D1.D2, factor 3 - numS = 1), mixing
// of both spins with the one step, and the potentials
// V_1 = alphaH L (d_1 + d_2) + U d_2, V_2 = alphaH L (d_1 + d_2) + U d_1
// (Main.f90:586-595).  Everything stays on the device; the host reads the
// window energies and a few scalars.

namespace forge {

// With A = a0 + a1, B = b0 + b1 and yX = L X: out[0] = A.yA, out[1] = A.yB,
// out[2] = B.yB, out[3] = a0.a1, out[4] = a0.b1 + b0.a1, out[5] = b0.b1
// (single block, fixed order; b*, yb* may be null: terms read 0).
__global__ void k_dots2(i64 n, const double* a0, const double* b0, const double* ya0,
                        const double* yb0, const double* a1, const double* b1, const double* ya1,
                        const double* yb1, double* out) {
  __shared__ double sh[6][256];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (i64 i = threadIdx.x; i < n; i += 256) {
    const double x0 = a0[i], x1 = a1[i];
    const double z0 = b0 ? b0[i] : 0.0, z1 = b1 ? b1[i] : 0.0;
    const double A = x0 + x1, B = z0 + z1;
    const double yA = ya0[i] + ya1[i], yB = (yb0 ? yb0[i] : 0.0) + (yb1 ? yb1[i] : 0.0);
    acc[0] += A * yA;
    acc[1] += A * yB;
    acc[2] += B * yB;
    acc[3] += x0 * x1;
    acc[4] += x0 * z1 + z0 * x1;
    acc[5] += z0 * z1;
  }
  for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 6; ++q) out[q] = sh[q][0];
}

// d_s = (1-l) da_s + l db_s, y = (1-l)(ya0 + ya1) + l (yb0 + yb1):
// pot0 = alphaH y + U d1, pot1 = alphaH y + U d0; da_s <- d_s (db*, yb* may
// be null when l = 0)
__global__ void k_potential2(i64 n, double* da0, const double* db0, const double* ya0,
                             const double* yb0, double* da1, const double* db1, const double* ya1,
                             const double* yb1, double l, double alphaH, double U, double* pot0,
                             double* pot1) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double u = 1.0 - l;
  double y = ya0[i] + ya1[i], d0 = da0[i], d1 = da1[i];
  if (db0) {
    y = u * y + l * (yb0[i] + yb1[i]);
    d0 = u * d0 + l * db0[i];
    d1 = u * d1 + l * db1[i];
  }
  pot0[i] = alphaH * y + U * d1;
  pot1[i] = alphaH * y + U * d0;
  da0[i] = d0;
  da1[i] = d1;
}

// Energies (Ekin_1 + Ekin_2, EF_1 + EF_2, EH, EU) of (1-l) a + l b from the
// per-spin pieces of energy_pieces and the cross-spin dots of k_dots2.
static void energies2_at(const double* pc0, const double* pc1, const double* x, double l, double* e) {
  const double u = 1.0 - l;
  e[0] = 0.0;
  e[1] = 0.0;
  for (const double* pc : {pc0, pc1}) {
    e[0] += u * (pc[0] + pc[11]) + l * (pc[1] + pc[12]);
    e[1] += u * u * pc[2] + 2.0 * u * l * pc[3] + l * l * pc[4];
  }
  e[2] = 0.5 * (u * u * x[0] + 2.0 * u * l * x[1] + l * l * x[2]);
  e[3] = u * u * x[3] + u * l * x[4] + l * l * x[5];
}

static double total_energy2(const forge_params& P, const double* e) {
  return P.alpha * e[1] + e[0] + P.U * e[3] + P.alphaH * e[2];
}

}  // namespace forge

struct forge_scf2 {
  forge_ctx* c[2] = {nullptr, nullptr};
  double* x = nullptr;  // device: the 6 cross-spin dots
  bool ready = false, mixed_done = false;
  double e[4] = {0, 0, 0, 0}, Etot = 0;
  int it = 0;
};

extern "C" int forge_scf2_create(forge_ctx* spin1, forge_ctx* spin2, forge_scf2** out) {
  if (!spin1 || !spin2 || !out || spin1 == spin2) {
    set_error("forge_scf2_create: two distinct contexts expected");
    return -2;
  }
  const forge_params &a = spin1->P, &b = spin2->P;
  if (a.ndim != b.ndim || a.nnc != b.nnc || a.numk != b.numk || a.nLower != b.nLower ||
      a.nUpper != b.nUpper || a.nlayers != b.nlayers || a.nfilling != b.nfilling || a.U != b.U ||
      a.alpha != b.alpha || a.alphaH != b.alphaH || a.Delta != b.Delta ||
      a.energy_tol != b.energy_tol || a.step_alt != b.step_alt || spin1->Coords != spin2->Coords) {
    set_error("forge_scf2_create: the two contexts must describe the same system");
    return -2;
  }
  auto* p = new forge_scf2();
  p->c[0] = spin1;
  p->c[1] = spin2;
  if (dalloc((void**)&p->x, sizeof(double) * 8)) {
    delete p;
    return -4;
  }
  *out = p;
  return 0;
}

extern "C" int forge_scf2_destroy(forge_scf2* p) {
  if (!p) return 0;
  if (p->x) (void)hipFree(p->x);
  delete p;
  return 0;
}

// Options, C3 set, symmetry and valley flags must agree between the spins;
// the two-spin loop follows the reference's T = 0 rule and OptimalStep.
static int scf2_check(forge_scf2* p, const char* who) {
  forge_ctx *c0 = p->c[0], *c1 = p->c[1];
  const forge_scf_options &o0 = c0->opt, &o1 = c1->opt;
  if (o0.occ_mode != o1.occ_mode || o0.mix_mode != o1.mix_mode || o0.precision != o1.precision ||
      o0.etol != o1.etol || o0.switch_de != o1.switch_de || c0->c3 != c1->c3 ||
      c0->valley != c1->valley || c0->sym != c1->sym) {
    set_error(std::string(who) + ": the two spins' options, C3 set, valley and symmetry flags differ");
    return -2;
  }
  if (o0.occ_mode == 2 || o0.mix_mode != 0) {
    set_error(std::string(who) + ": two spins follow the reference's T = 0 rule (occ_mode 0 or 1) "
              "and OptimalStep (mix_mode 0)");
    return -2;
  }
  if (c0->valley) {
    set_error(std::string(who) + ": the valley projection runs with one spin");
    return -2;
  }
  if (o0.graph >= 2 || o1.graph >= 2 || o0.timing || o1.timing) {
    set_error(std::string(who) + ": two spins run with graph 0 or 1 and without timing");
    return -2;
  }
  return 0;
}

static int scf2_solve(forge_ctx* c, const z2* rho, std::vector<double>& band) {
  return c->c3 ? solve_device_c3(c, rho, c->pot, band, false)
               : solve_device(c, rho, c->pot, band, false);
}

// energies and potentials of the spins' states c[s]->rho (Main.f90:290-346)
static int scf2_init_state(forge_scf2* p) {
  forge_ctx *c0 = p->c[0], *c1 = p->c[1];
  const i64 n = c0->n;
  double pc[2][13];
  for (int s = 0; s < 2; ++s) {
    CK(site_density(n, p->c[s]->rho, p->c[s]->da, p->c[s]->s));
    CK(energy_pieces(p->c[s], p->c[s]->rho, nullptr, pc[s]));  // synchronises c[s]->s
  }
  hipLaunchKernelGGL(k_dots2, dim3(1), dim3(256), 0, c0->s, n, c0->da, nullptr, c0->ya, nullptr,
                     c1->da, nullptr, c1->ya, nullptr, p->x);
  FORGE_LAUNCH();
  hipLaunchKernelGGL(k_potential2, dim3(cdiv(n, 256)), dim3(256), 0, c0->s, n, c0->da, nullptr,
                     c0->ya, nullptr, c1->da, nullptr, c1->ya, nullptr, 0.0, c0->P.alphaH, c0->P.U,
                     c0->pot, c1->pot);
  FORGE_LAUNCH();
  double x[6];
  FORGE_HIP(hipMemcpyAsync(x, p->x, sizeof(x), hipMemcpyDeviceToHost, c0->s));
  FORGE_HIP(hipStreamSynchronize(c0->s));
  energies2_at(pc[0], pc[1], x, 0.0, p->e);
  p->Etot = total_energy2(c0->P, p->e);
  p->it = 0;
  p->mixed_done = false;
  p->ready = true;
  for (forge_ctx* c : p->c) {
    c->scf_ready = false;  // the contexts' own loop does not continue the pair's state
    c->mixed_done = false;
    std::fill(c->tm, c->tm + FORGE_NTIMINGS, 0.0);
  }
  return 0;
}

extern "C" int forge_scf2_init(forge_scf2* p, const forge_z* rho_start) {
  if (!p) { set_error("forge_scf2_init: null pair"); return -2; }
  CK(scf2_check(p, "forge_scf2_init"));
  forge_ctx* c0 = p->c[0];
  const i64 n = c0->n, nnc = c0->nnc;
  const size_t tot = (size_t)nnc * n * n;
  int fails = 0;
  if (rho_start) {
    for (int s = 0; s < 2; ++s)
      FORGE_HIP(hipMemcpyAsync(p->c[s]->rho, (const z2*)rho_start + tot * s, sizeof(z2) * tot,
                               hipMemcpyHostToDevice, p->c[s]->s));
    for (int s = 0; s < 2; ++s) FORGE_HIP(hipStreamSynchronize(p->c[s]->s));
  } else {
    // first diagonalisation of spin 1 with rho = 0 and Potential = 0, its
    // density copied to spin 2 (Main.f90:238-288)
    FORGE_HIP(hipMemsetAsync(c0->pot, 0, sizeof(double) * n, c0->s));
    std::vector<double> band;
    fails = scf2_solve(c0, nullptr, band);
    if (fails < 0) return fails;
    Occupation occ;
    CK(window_step(c0, band, occ, false));
    FORGE_HIP(hipMemcpyAsync(c0->rho, c0->rho_new, sizeof(z2) * tot, hipMemcpyDeviceToDevice, c0->s));
    FORGE_HIP(hipMemcpyAsync(p->c[1]->rho, c0->rho_new, sizeof(z2) * tot, hipMemcpyDeviceToDevice, c0->s));
    FORGE_HIP(hipStreamSynchronize(c0->s));
  }
  CK(scf2_init_state(p));
  return fails;
}

// SortEnergies_2Spins (HartreeFock.f90:929-1004) over the two spins' sorted
// window energies key[0], key[1] (N each), with the window
// W = [[NP-1, NP-1], [NP+3, NP+3]] of Main.f90:418-421: lab[q] = 0 or 1.
// The reference fills its spin-2 lead-in from row ncount1 of spin 2 while
// advancing ncount2 (HartreeFock.f90:951-956); that is kept.  Its merge
// loops test the row bound after reading the row; the bound is tested first
// here (it decides either way).
static void sort_two_spins(const std::vector<double>* key, i64 N, i64 Nk, i64 NP,
                           std::vector<double>& E, std::vector<int>& lab) {
  const i64 w1 = NP - 1, w2 = NP + 3;
  E.clear();
  lab.clear();
  i64 c1 = 1, c2 = 1;
  auto put = [&](double v, int l) { E.push_back(v); lab.push_back(l); };
  auto e1 = [&](i64 q) { return key[0][q - 1]; };
  auto e2 = [&](i64 q) { return key[1][q - 1]; };
  for (i64 q = 1; q <= (w1 - 1) * Nk; ++q) { put(e1(c1), 0); ++c1; }
  for (i64 q = 1; q <= (w1 - 1) * Nk; ++q) { put(e2(c1), 1); ++c2; }
  while (c1 <= Nk * (w2 - 1) && c2 <= Nk * (w2 - 1)) {
    if (e1(c1) <= e2(c2)) {
      put(e1(c1), 0); ++c1;
      while (c1 <= N && e1(c1) <= e2(c2)) { put(e1(c1), 0); ++c1; }
      put(e2(c2), 1); ++c2;
    } else {
      put(e2(c2), 1); ++c2;
      while (c2 <= N && e2(c2) <= e1(c1)) { put(e2(c2), 1); ++c2; }
      put(e1(c1), 0); ++c1;
    }
  }
  while (c1 <= N) { put(e1(c1), 0); ++c1; }
  while (c2 <= N) { put(e2(c2), 1); ++c2; }
}

// One two-spin SCF iteration; traj_row (12 values) may be null.
static int scf2_step(forge_scf2* p, double* traj_row, bool* stop) {
  forge_ctx *c0 = p->c[0], *c1 = p->c[1];
  const i64 n = c0->n, nnc = c0->nnc, Nk = c0->Nk, N = c0->nb * Nk;
  const size_t tot = (size_t)nnc * n * n;
  *stop = false;
  const bool mixed = c0->opt.precision == 1 && !p->mixed_done;
  for (forge_ctx* c : p->c) {
    c->mixed_done = p->mixed_done;
    if (mixed) c->tm[12] += 1;
  }
  int fails = 0;
  std::vector<double> band[2], key[2], val[2];
  for (int s = 0; s < 2; ++s) {
    const int f = scf2_solve(p->c[s], p->c[s]->rho, band[s]);
    if (f < 0) return f;
    fails += f;
    sort_window(p->c[s], band[s], key[s], val[s]);
  }
  // two-spin Fermi block (Main.f90:416-458)
  const i64 NP = (n / 2 - c0->nL) + 1;
  std::vector<double> E;
  std::vector<int> lab;
  sort_two_spins(key, N, Nk, NP, E, lab);
  const double xf = (double)(Nk * NP) + (double)Nk / 2.0 * c0->P.nfilling;
  const i64 nF = 2 * (i64)(xf >= 0 ? std::floor(xf + 0.5) : -std::floor(-xf + 0.5));
  if (nF < 1 || nF >= 2 * N) {
    set_error("forge_scf2_run: the Fermi level lies outside the two spins' window");
    return -2;
  }
  const double EF = (E[nF - 1] + E[nF]) / 2.0;
  i64 nOcc[2] = {0, 0}, nPart[2] = {0, 0};
  for (i64 q = 0; q < 2 * N; ++q) {
    const int sp = lab[q];
    if ((E[q] - EF) < -c0->P.energy_tol) ++nOcc[sp];
    else if (std::fabs(E[q] - EF) < c0->P.energy_tol) { ++nPart[sp]; ++nOcc[sp]; }
  }
  const double Deg = (nPart[0] != 0 || nPart[1] != 0)
                         ? 1.0 / (double)(nPart[0] + nPart[1]) *
                               (double)(nF + nPart[0] + nPart[1] - nOcc[0] - nOcc[1])
                         : 0.0;
  // GetFock per spin with its own counts and the shared DegFactor, then the
  // symmetry enforcement of each spin's new density (Main.f90:514-534)
  for (int s = 0; s < 2; ++s) {
    forge_ctx* c = p->c[s];
    std::vector<double> w, wi;
    zero_temperature_weights(c->opt.occ_mode, N, nOcc[s], nPart[s], Deg, w, wi);
    Occupation occ;
    weighted_states(c, val[s], w, wi, occ);
    CK(window_of(c, occ, true));
    if (c->sym) CK(symmetrize(n, (int)nnc, c->rho_new, c->rho_new + tot, c->sym, c->c2_perm, c->s));
  }
  // OptimalStep on the two-spin energy
  double pc[2][13];
  for (int s = 0; s < 2; ++s) {
    CK(site_density(n, p->c[s]->rho_new, p->c[s]->db, p->c[s]->s));
    CK(energy_pieces(p->c[s], p->c[s]->rho, p->c[s]->rho_new, pc[s]));  // synchronises c[s]->s
  }
  hipLaunchKernelGGL(k_dots2, dim3(1), dim3(256), 0, c0->s, n, c0->da, c0->db, c0->ya, c0->yb,
                     c1->da, c1->db, c1->ya, c1->yb, p->x);
  FORGE_LAUNCH();
  double x[6];
  FORGE_HIP(hipMemcpyAsync(x, p->x, sizeof(x), hipMemcpyDeviceToHost, c0->s));
  std::vector<double> hda[2], hdb[2];
  for (int s = 0; s < 2; ++s) {
    hda[s].resize(n);
    hdb[s].resize(n);
    FORGE_HIP(hipMemcpyAsync(hda[s].data(), p->c[s]->da, sizeof(double) * n, hipMemcpyDeviceToHost, c0->s));
    FORGE_HIP(hipMemcpyAsync(hdb[s].data(), p->c[s]->db, sizeof(double) * n, hipMemcpyDeviceToHost, c0->s));
  }
  FORGE_HIP(hipStreamSynchronize(c0->s));
  const double ein[4] = {p->e[0], p->e[1], p->e[2], p->e[3]};
  const double EtotIn = p->Etot;
  double e1[4], eh[4];
  energies2_at(pc[0], pc[1], x, 1.0, e1);
  energies2_at(pc[0], pc[1], x, 0.5, eh);
  const double step = scf_oda_step(EtotIn, total_energy2(c0->P, e1), total_energy2(c0->P, eh),
                                   c0->P.step_alt);
  energies2_at(pc[0], pc[1], x, step, p->e);
  p->Etot = total_energy2(c0->P, p->e);
  // mixing of both spins with the one step, then the potentials
  for (int s = 0; s < 2; ++s) CK(mix(tot, p->c[s]->rho, p->c[s]->rho_new, step, p->c[s]->s));
  hipLaunchKernelGGL(k_potential2, dim3(cdiv(n, 256)), dim3(256), 0, c0->s, n, c0->da, c0->db,
                     c0->ya, c0->yb, c1->da, c1->db, c1->ya, c1->yb, step, c0->P.alphaH, c0->P.U,
                     c0->pot, c1->pot);
  FORGE_LAUNCH();
  for (int s = 0; s < 2; ++s) FORGE_HIP(hipStreamSynchronize(p->c[s]->s));
  // DensityConvergence: max over spins and sites |D_mixed - D_in| (Main.f90:564-571)
  double dconv = 0;
  for (int s = 0; s < 2; ++s)
    for (i64 q = 0; q < n; ++q) {
      const double dmix = (1.0 - step) * hda[s][q] + step * hdb[s][q];
      dconv = std::max(dconv, std::fabs(dmix - hda[s][q]));
    }
  ++p->it;
  if (traj_row) {
    double* r = traj_row;
    r[0] = (double)p->it; r[1] = EtotIn; r[2] = ein[0]; r[3] = ein[1]; r[4] = ein[2];
    r[5] = ein[3]; r[6] = dconv; r[7] = step; r[8] = EtotIn - p->Etot; r[9] = EF;
    r[10] = (double)(nOcc[0] + nOcc[1]); r[11] = (double)(nPart[0] + nPart[1]);
  }
  const double dE = std::fabs(EtotIn - p->Etot);
  if (mixed) {
    const double sw = c0->opt.switch_de > 0 ? c0->opt.switch_de : 1e-4;
    if (dE < sw || (c0->opt.etol > 0 && dE < c0->opt.etol)) p->mixed_done = true;
    return fails;
  }
  if (c0->opt.etol > 0 && dE < c0->opt.etol) *stop = true;
  return fails;
}

extern "C" int forge_scf2_run(forge_scf2* p, int64_t max_iter, double* traj, int64_t* iters_done) {
  if (!p || max_iter < 0) { set_error("forge_scf2_run: invalid argument"); return -2; }
  if (!p->ready) { set_error("forge_scf2_run: call forge_scf2_init first"); return -5; }
  CK(scf2_check(p, "forge_scf2_run"));
  int fails = 0;
  i64 done = 0;
  for (i64 itn = 0; itn < max_iter; ++itn) {
    bool stop = false;
    const int f = scf2_step(p, traj ? traj + 12 * itn : nullptr, &stop);
    if (f < 0) return f;
    fails += f;
    ++done;
    if (stop) break;
  }
  if (iters_done) *iters_done = done;
  return fails;
}

extern "C" int forge_scf2_get_density(forge_scf2* p, forge_z* rho) {
  if (!p || !rho) { set_error("forge_scf2_get_density: null argument"); return -2; }
  const size_t tot = (size_t)p->c[0]->nnc * p->c[0]->n * p->c[0]->n;
  for (int s = 0; s < 2; ++s) {
    FORGE_HIP(hipMemcpyAsync((z2*)rho + tot * s, p->c[s]->rho, sizeof(z2) * tot, hipMemcpyDeviceToHost,
                             p->c[s]->s));
    FORGE_HIP(hipStreamSynchronize(p->c[s]->s));
  }
  return 0;
}

// ---------------------------------------------------------------- twist-angle sweeps
// Several systems of different orders (e.g. TBG cells m = 8..14, each with its
// own filling and k grid) advance their SCF iterations together.  Per
// iteration all their H(k) go into ONE eigensolver batch: every matrix gets an
// Nmax x Nmax slot (Nmax the largest order) and is assembled into the slot's
// trailing n_i x n_i corner; the slots are ordered by decreasing n_i, so the
// blocked Householder reduction runs one set of launches per column over the
// prefix of matrices that have started (first_col = Nmax - n_i), all with the
// same trailing order, and spends no work on the padding.  The tridiagonal
// eigensolver and back-transformation then run per system (batched over its
// k-points) on the compacted d, e, tau, and each system finishes its
// iteration (densities, Fermi level, energies, mixing) on its own buffers.
// A system that has converged leaves the batch.
struct forge_sweep {
  std::vector<forge_ctx*> sys;
  i64 Nmax = 0, ldH = 0;
  int Ktot = 0;
  z2* H = nullptr;
  void* trd_ws = nullptr;
  TridiagWork trd{};
  hipStream_t s = nullptr;
  std::vector<i64> first_col;
  std::vector<std::pair<int, int>> assigned;  // (system, first slot) of the last iteration
  float2* S = nullptr;  // FP32 copy of the slots (allocated when a system asks for precision 1)
};

// zero columns [0, c1) of K consecutive slots (all Nmax rows)
__global__ void k_sweep_clear_cols(i64 nrow, i64 c1, z2* H, i64 ld, size_t sH) {
  z2* h = H + blockIdx.z * sH;
  for (i64 c = blockIdx.y; c < c1; c += gridDim.y)
    for (i64 r = (i64)blockIdx.x * 256 + threadIdx.x; r < nrow; r += (i64)gridDim.x * 256)
      h[r + ld * c] = mk(0.0, 0.0);
}

// d, e, tau of system-local columns 0..ni-1 (sweep columns s0..) of K slots
__global__ void k_sweep_gather(i64 Nmax, i64 s0, i64 ni, const double* __restrict__ d,
                               const double* __restrict__ e, const z2* __restrict__ tau,
                               double* __restrict__ di, double* __restrict__ ei,
                               z2* __restrict__ ti) {
  const int k = blockIdx.y;
  for (i64 t = (i64)blockIdx.x * 256 + threadIdx.x; t < ni; t += (i64)gridDim.x * 256) {
    const size_t src = (size_t)k * Nmax + s0 + t, dst = (size_t)k * ni + t;
    di[dst] = d[src];
    ei[dst] = t < ni - 1 ? e[src] : 0.0;
    ti[dst] = t < ni - 1 ? tau[src] : mk(0.0, 0.0);
  }
}

extern "C" int forge_sweep_destroy(forge_sweep* w);

extern "C" int forge_sweep_create(forge_ctx* const* ctxs, int64_t nsys, forge_sweep** out) {
  if (!ctxs || nsys < 1 || !out) { set_error("forge_sweep_create: invalid argument"); return -2; }
  *out = nullptr;
  auto* w = new forge_sweep();
  for (i64 i = 0; i < nsys; ++i) {
    forge_ctx* c = ctxs[i];
    if (!c) { delete w; set_error("forge_sweep_create: null context"); return -2; }
    if (c->c3) {
      delete w;
      set_error("forge_sweep_create: the C3-reduced k-set is not batched in sweeps");
      return -2;
    }
    if (c->Kb < c->Nk) {
      delete w;
      set_error("forge_sweep_create: every system must solve all its k-points in one batch");
      return -4;
    }
    w->sys.push_back(c);
    w->Nmax = std::max<i64>(w->Nmax, c->n);
    w->Ktot += (int)c->Nk;
  }
  w->ldH = pad_ld(w->Nmax);
  const size_t sH = (size_t)w->ldH * w->Nmax;
  const hipError_t se = hipStreamCreateWithFlags(&w->s, hipStreamNonBlocking);
  if (se != hipSuccess) {
    set_error(std::string("forge_sweep_create: hipStreamCreateWithFlags: ") + hipGetErrorString(se));
    w->s = nullptr;
    forge_sweep_destroy(w);
    return -1;
  }
  if (dalloc((void**)&w->H, sizeof(z2) * sH * w->Ktot) ||
      dalloc(&w->trd_ws, tridiag_workspace_bytes(w->Nmax, w->Ktot))) {
    forge_sweep_destroy(w);
    return -4;
  }
  tridiag_bind(w->trd, w->trd_ws, w->Nmax, w->Ktot);
  w->trd.A = w->H;
  w->trd.sA = sH;
  w->trd.lda = w->ldH;
  const hipError_t me = hipMemsetAsync(w->H, 0, sizeof(z2) * sH * w->Ktot, w->s);
  const hipError_t ye = me == hipSuccess ? hipStreamSynchronize(w->s) : me;
  if (ye != hipSuccess) {
    set_error(std::string("forge_sweep_create: clearing the slots: ") + hipGetErrorString(ye));
    forge_sweep_destroy(w);
    return -1;
  }
  *out = w;
  return 0;
}

extern "C" int forge_sweep_destroy(forge_sweep* w) {
  if (!w) return 0;
  if (w->s) (void)hipStreamSynchronize(w->s);
  tridiag_release(w->trd);
  if (w->H) (void)hipFree(w->H);
  if (w->S) (void)hipFree(w->S);
  if (w->trd_ws) (void)hipFree(w->trd_ws);
  if (w->s) (void)hipStreamDestroy(w->s);
  delete w;
  return 0;
}

// the batched eigensolve of one sweep iteration over the systems in `act`
static int sweep_solve(forge_sweep* w, const std::vector<int>& act) {
  const size_t sH = (size_t)w->ldH * w->Nmax;
  std::vector<std::pair<int, int>> assign;
  w->first_col.clear();
  int slot = 0;
  for (int i : act) {
    forge_ctx* c = w->sys[i];
    assign.push_back({i, slot});
    for (i64 k = 0; k < c->Nk; ++k) w->first_col.push_back(w->Nmax - c->n);
    slot += (int)c->Nk;
  }
  // a slot's columns left of its matrix must be zero; they change only when
  // the assignment of systems to slots does
  if (assign != w->assigned) {
    for (auto& a : assign) {
      forge_ctx* c = w->sys[a.first];
      const i64 s0 = w->Nmax - c->n;
      if (s0 == 0) continue;
      hipLaunchKernelGGL(k_sweep_clear_cols, dim3((unsigned)cdiv(w->Nmax, 256),
                                                  (unsigned)std::min<i64>(s0, 1024), (unsigned)c->Nk),
                         dim3(256), 0, w->s, w->Nmax, s0, w->H + a.second * sH, w->ldH, sH);
      FORGE_LAUNCH();
    }
    w->assigned = assign;
  }
  for (auto& a : assign) {
    forge_ctx* c = w->sys[a.first];
    const i64 s0 = w->Nmax - c->n;
    CK(assemble(c->n, (int)c->nnc, (int)c->Nk, c->rho, c->V, c->T, c->pot, c->bias, c->P.alpha,
                c->ph_asm, w->H + a.second * sH + s0 + w->ldH * s0, w->ldH, sH, w->s));
  }
  // the shared reduction streams the FP32 copy while every system in the
  // batch is in its mixed phase; each system's back-transform follows its own
  bool mixed_all = w->S != nullptr;
  for (int i : act) mixed_all = mixed_all && mixed_phase(w->sys[i]);
  w->trd.K = slot;
  w->trd.first_col = w->first_col.data();
  w->trd.S = mixed_all ? w->S : nullptr;
  w->trd.sS = sH;
  w->trd.use_graph = 0;
  for (int i : act) w->trd.use_graph |= w->sys[i]->opt.graph;
  CK(tridiag(w->trd, w->s, nullptr, nullptr, nullptr));
  int fails = 0;
  for (auto& a : assign) {
    forge_ctx* c = w->sys[a.first];
    const i64 n = c->n, s0 = w->Nmax - n;
    const int K = (int)c->Nk;
    const size_t q0 = (size_t)a.second;
    hipLaunchKernelGGL(k_sweep_gather, dim3((unsigned)cdiv(n, 256), K), dim3(256), 0, w->s, w->Nmax,
                       s0, n, w->trd.d + q0 * w->Nmax, w->trd.e + q0 * w->Nmax,
                       w->trd.tau + q0 * w->Nmax, c->trd.d, c->trd.e, c->trd.tau);
    FORGE_LAUNCH();
    int f = dc_solve(c->trd.d, c->trd.e, n, K, c->nU, c->evals, c->Zq, c->dc_ws, w->s);
    if (f < 0) return f;
    fails += f;
    // reflectors of this system: its slots' trailing corners, compacted tau;
    // the batch's reduction is one-stage, whatever the system's own last
    // reduction was (its two-stage flags would apply stale chase reflectors)
    TridiagWork v = c->trd;
    v.ts = nullptr;
    v.two_stage = w->trd.two_stage;
    v.voff = w->trd.voff;
    v.n = n;
    v.K = K;
    v.A = w->H + q0 * sH + s0 + w->ldH * s0;
    v.sA = sH;
    v.lda = w->ldH;
    CK(backtransform(v, c->Zq, n, (size_t)n * c->nU, c->nU, c->Cv, n, (size_t)n * c->nU, c->bt_ws,
                     w->s, mixed_phase(c) ? c->S : nullptr));
  }
  return fails;
}

extern "C" int forge_sweep_run(forge_sweep* w, int64_t max_iter, double* traj, int64_t* iters_done) {
  if (!w || max_iter < 0) { set_error("forge_sweep_run: invalid argument"); return -2; }
  const int ns = (int)w->sys.size();
  bool any_mixed = false;
  for (forge_ctx* c : w->sys) {
    if (!c->scf_ready) { set_error("forge_sweep_run: call forge_scf_init on every system"); return -5; }
    any_mixed = any_mixed || c->opt.precision == 1;
    FORGE_HIP(hipStreamSynchronize(c->s));
  }
  if (any_mixed && !w->S) {
    const size_t bytes = sizeof(float2) * (size_t)w->ldH * w->Nmax * w->Ktot;
    size_t fr = 0, tot = 0;
    FORGE_HIP(hipMemGetInfo(&fr, &tot));
    if (bytes + ((size_t)256 << 20) > fr) {
      set_error("forge_sweep_run: not enough device memory for the FP32 copy of precision 1");
      return -4;
    }
    CK(dalloc((void**)&w->S, bytes));
  }
  // every system's work goes to the sweep's stream for the run
  std::vector<hipStream_t> own(ns);
  for (int i = 0; i < ns; ++i) {
    own[i] = w->sys[i]->s;
    w->sys[i]->s = w->s;
    std::fill(w->sys[i]->tm, w->sys[i]->tm + FORGE_NTIMINGS, 0.0);
  }
  std::vector<i64> done(ns, 0);
  std::vector<bool> stopped(ns, false);
  int fails = 0, rc = 0;
  for (i64 itn = 0; itn < max_iter && rc >= 0; ++itn) {
    std::vector<int> act;
    for (int i = 0; i < ns; ++i)
      if (!stopped[i]) act.push_back(i);
    if (act.empty()) break;
    std::stable_sort(act.begin(), act.end(),
                     [&](int a, int b) { return w->sys[a]->n > w->sys[b]->n; });
    int f = sweep_solve(w, act);
    if (f < 0) { rc = f; break; }
    fails += f;
    for (int i : act) {
      bool stop = false;
      forge_ctx* c = w->sys[i];
      f = scf_step(c, traj ? traj + ((size_t)i * max_iter + itn) * 12 : nullptr, false, true, &stop);
      if (f < 0) { rc = f; break; }
      fails += f;
      ++done[i];
      if (stop) stopped[i] = true;
    }
  }
  for (int i = 0; i < ns; ++i) {
    w->sys[i]->s = own[i];
    if (iters_done) iters_done[i] = done[i];
  }
  FORGE_HIP(hipStreamSynchronize(w->s));
  return rc < 0 ? rc : fails;
}
