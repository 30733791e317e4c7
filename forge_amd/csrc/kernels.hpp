// Internal kernel entry points (host wrappers) of libforge_hip.
#pragma once
#include <vector>

#include "common.hpp"

namespace forge {

constexpr int MAXCELL = 8;

struct CellShifts {
  double n1[MAXCELL], n2[MAXCELL];
  double t1x, t1y, t2x, t2y;
};

struct DeviceGeom {
  i64 n;
  int nnc;
  const double *X, *Y, *Z;
  CellShifts cells;
};

// Energy model of OptimalStep (HartreeFock.f90:1148-1227) from the pieces of
// one energy pass over a = rho_in and b = rho_new: pc[0..4] = K_a, K_b, F_aa,
// F_ab, F_bb; [5..7] = H_aa, H_ab, H_bb (d.L.d); [8..10] = U_aa, U_ab, U_bb;
// [11..12] bias terms of a, b.  Energies (Ekin, EF, EH, EU) of (1-l) a + l b
// -- the energy is quadratic in rho, so these are exact for every l -- and
// the total energy (Main.f90:312,556), shared by the host and device loops.
__host__ __device__ inline void scf_energies_at(const double* pc, double l, double* e) {
  const double u = 1.0 - l;
  e[0] = u * (pc[0] + pc[11]) + l * (pc[1] + pc[12]);
  e[1] = u * u * pc[2] + 2.0 * u * l * pc[3] + l * l * pc[4];
  e[2] = 2.0 * (u * u * pc[5] + 2.0 * u * l * pc[6] + l * l * pc[7]);
  e[3] = u * u * pc[8] + 2.0 * u * l * pc[9] + l * l * pc[10];
}
__host__ __device__ inline double scf_total_energy(double alpha, double alphaH, double U,
                                                   const double* e) {
  return alpha * e[1] * 2.0 + e[0] * 2.0 + U * e[3] + alphaH * e[2];
}
// OptimalStep's choice (HartreeFock.f90:1162-1208) from E(0) = c0, E(1), E(1/2)
__host__ __device__ inline double scf_oda_step(double c0, double E1, double Eh, double step_alt) {
  const double c1 = 4.0 * Eh - E1 - 3.0 * c0, c2 = E1 - c0 - c1;
  if (c2 > 0 && fabs(-c1 / 2.0 / c2 - 0.5) < 0.5) return -c1 / 2.0 / c2;
  if (E1 < c0) return 1.0;
  return step_alt;
}

// operators.hip
int pair_tables(const DeviceGeom& g, const PairParams& pp, double* V, double* T, hipStream_t s);
int longrange(const DeviceGeom& g, const PairParams& pp, int numC, double* L, hipStream_t s);
int assemble(i64 n, int nnc, int K, const z2* rho, const double* V, const double* T,
             const double* pot, const double* bias, double alpha, const double* phase, z2* H,
             size_t ldH, size_t strideH, hipStream_t s);
int energy_pass(size_t nn, int nnc, const z2* a, const z2* b, const double* V, const double* T,
                double* part, int nblocks, double* out5, hipStream_t s);
int hartree_matvec(i64 n, const double* L, const double* da, const double* db, double* ya,
                   double* yb, hipStream_t s);
int mix(size_t tot, z2* a, const z2* b, double s, hipStream_t st);
int site_density(i64 n, const z2* rho, double* d, hipStream_t s);

// tridiag.hip: batched Hermitian -> real symmetric tridiagonal (lower storage).
struct TridiagWork {
  i64 n;
  int K;
  z2* A;          // K matrices, column-major, leading dimension lda, stride sA
  size_t sA;
  i64 lda;        // >= n; a multiple of 16 keeps 16-row column chunks 256-byte aligned
  double* d;      // K x n
  double* e;      // K x n
  z2* tau;        // K x n
  z2* W;          // K x n x NB
  z2* P;          // hemv row partials: K x nt x n   (A v over column strip J, row r)
  z2* P2;         // hemv column partials: K x nt x nseg x TS (A^H v of strip J, segment s)
  z2* tpart;      // K x nseg x PB x 2NB: W^H v, V^H v partials per row chunk
  double* npart;  // K x nblk: column-norm partials of k_trd_col
  z2* dpart;      // K x nblk: w^H v partials of k_trd_fin
  z2* scal;       // K x 4: tau, 1/(alpha - beta), beta of the current column
  float2* S;      // null, or an FP32 copy of A (same lda, stride sS) that the hemv
  size_t sS;      // streams instead of A; k_trd_her2k keeps it current
  hipEvent_t* evs;  // event pairs of the sampled hemv launches (created lazily)
  i64 nev;
  // Uneven sizes (null: all K matrices are n x n).  Matrix k holds an
  // n_k x n_k matrix in the trailing corner of its n x n slot (first_col[k] =
  // n - n_k, non-decreasing in k, the rest of the slot's columns zero); it is
  // reduced from column first_col[k] on, so at column j the matrices with
  // first_col <= j are a prefix of the batch and share the trailing order.
  const i64* first_col;  // host array of K entries
  // nonzero: tridiag() captures its launch sequence into a HIP graph on the
  // first call for a shape (n, K, first_col, FP32 copy or not, buffers) and
  // replays it while the shape is unchanged (untimed calls only)
  int use_graph;
  void* graph;  // the captured graph and its shape; tridiag_release frees it
  // Two-stage reduction (twostage.hip): non-null ts and no first_col make
  // tridiag() reduce to a band of width TS_KD first and chase the band to
  // tridiagonal form.  The reflectors of the first stage then have their
  // unit element TS_KD rows below their column (voff = TS_KD; the one-stage
  // reflectors: voff = 1), and the back-transformation applies the chase's
  // reflectors (ts) before them.  Set by every reduction.
  struct TwoStageWork* ts;
  int voff;
  int two_stage;
};
constexpr int TRD_NB = 32;  // panel width (a multiple of 8, at most 128; 64 measured slower)
constexpr int TRD_TS = 32;     // hemv column-strip width
constexpr int TRD_SEGR = 1024;   // the smallest hemv segment height (sizes the partial arrays);
                                 // the height per trailing order is tuned (1024 or 2048)
constexpr int TRD_PB = 4;      // hemv panel-product blocks per row segment
size_t tridiag_workspace_bytes(i64 n, int K);
int tridiag_bind(TridiagWork& w, void* base, i64 n, int K);
// the hemv variants of w's shape per band of trailing orders (bands start at
// band_lo[0..3]): code = segment height / 1024 + 2 x lane layout (1, 2: 16
// lanes per column chunk, 1024 or 2048 rows; 3, 4: 8 lanes with 16-byte
// row-pair loads); returns 1 when tuned, 0 for the defaults
int tridiag_hemv_plan(const TridiagWork& w, int* code_by_band, i64* band_lo);
int tridiag(TridiagWork& w, hipStream_t s, double* hemv_ms, long* hemv_launches,
            double* hemv_bytes);
// the reduction's launches alone, for an enclosing stream capture (no host
// synchronisation, no graph of its own)
int tridiag_enqueue(TridiagWork& w, hipStream_t s);
void tridiag_release(TridiagWork& w);

// twostage.hip: the two-stage reduction.  Stage 1 (dense -> band of width
// TS_KD): per panel of TS_KD columns a Householder QR of the panel below its
// diagonal block, then A22 <- Q^H A22 Q as one matrix-core product of the
// trailing matrix with the panel's reflectors (A22 V) and a rank-2 TS_KD
// update.  Stage 2 (band -> tridiagonal): bulge chasing in band storage, one
// launch per slot of independent steps.  Its reflectors are kept in groups
// of TS_NB consecutive sweeps for the back-transformation.
constexpr int TS_KD = 32;             // band width = panel width
constexpr int TS_NB = 32;             // sweeps per group of the chase's back-transformation
constexpr int TS_LDAB = 2 * TS_KD + 1;  // band storage: column c holds A(c .. c + 2 TS_KD, c)
constexpr int TS_YSPLIT = 8;          // maximum column splits of the panel product A22 V
struct TwoStageWork {
  i64 n = 0;
  int K = 0;
  z2* V = nullptr;      // K x n x KD: the panel's reflectors (unit diagonal, zeros above), rows global
  z2* X = nullptr;      // K x n x KD: A22 V T
  z2* Y = nullptr;      // K x YSPLIT x n x KD: partial A22 V per column split
  z2* qpart = nullptr;  // K x nbq x KD: v^H A partials of the panel QR
  z2* gpart = nullptr;  // K x nbq x KD x KD: V^H V partials
  double* npart = nullptr;  // K x nbq: column-norm partials
  z2* T = nullptr;      // K x KD x KD: the panel's triangular factor
  z2* Zm = nullptr;     // K x KD x KD: T^H V^H X
  z2* mpart = nullptr;  // K x nbx x KD x KD: V^H X partials
  z2* AB = nullptr;     // K x n x TS_LDAB: the band
  z2* V2 = nullptr;     // K x nblk2 x NB x KD: chase reflectors by (group, step)
  z2* T2 = nullptr;     // K x nblk2 x NB x NB: triangular factor of each block
  z2* tau2 = nullptr;   // K x nblk2 x NB
  i64* goff = nullptr;  // device: first block of each group (G + 1 entries)
  std::vector<i64> hoff;  // host copy
  i64 G = 0, nblk2 = 0;
  // Timing of the chase's back-transformation (k_ts_q2): while q2_timed is
  // set, every TS_Q2_STRIDE-th group launch is bracketed by an event pair on
  // the dispatch; twostage_q2_timing reads and clears the records.
  bool q2_timed = false;
  mutable std::vector<hipEvent_t> q2ev;
  mutable std::vector<double> q2flops;  // algorithmic flops of each recorded launch
  mutable std::vector<double> q2bytes;  // its eigenvector rows' bytes, read and written once
  mutable size_t q2rec = 0;
};
constexpr int TS_Q2_STRIDE = 4;
size_t twostage_workspace_bytes(i64 n, int K);
int twostage_bind(TwoStageWork& ts, void* base, i64 n, int K);
// the reduction's launches (no host synchronisation while capturing)
int twostage_reduce(TridiagWork& w, hipStream_t s, bool capturing);
// its trailing update on the one-stage her2k kernel (tridiag.hip)
int her2k_twostage(TridiagWork& w, i64 p, i64 i0, int nbp, hipStream_t s);
// C <- Q2 C (the chase's reflectors), C complex FP64 or FP32
int twostage_apply_q2(const TridiagWork& t, i64 m, z2* C, size_t ldC, size_t sC, hipStream_t s);
int twostage_apply_q2(const TridiagWork& t, i64 m, float2* C, size_t ldC, size_t sC, hipStream_t s);
// the recorded k_ts_q2 launches since the last call: count, summed time (ms),
// algorithmic flops (16 per reflector element per column: v^H c and
// c -= tau v w as complex multiply-adds) and the bytes of the eigenvector
// rows they pass over (read and written once); waits for the last one
int twostage_q2_timing(const TwoStageWork& ts, long* launches, double* ms, double* flops,
                       double* bytes);
void twostage_release_events(TwoStageWork& ts);

// dc.hip: batched divide-and-conquer for real symmetric tridiagonals.
struct DCWork;
size_t dc_workspace_bytes(i64 n, int K, i64 mout);
// deferred: no host synchronisation (graph capture): workspace clears by
// kernel, the failure count stays on the device (dc_fail_ptr) instead of
// being returned, and the merge descriptors are the ones a non-deferred call
// with the same workspace, n, K and mout left there
int dc_solve(const double* d, const double* e, i64 n, int K, i64 mout, double* evals,
             double* Q, void* work, hipStream_t s, bool deferred = false);
const int* dc_fail_ptr(void* work, i64 n, int K);

// backtr.hip: C(n x m, complex) = Q_house * Z(n x m, real), Q from tridiag.
// C32 non-null: the products run in complex FP32 in C32 (n x m per matrix,
// ld n), and C receives the columns normalised in FP64.  The block width of
// the compact WY products is tuned per shape on first use (*width_used);
// fixed_width (64, 128 or 256; FP64 only) skips the tuning.
size_t backtr_workspace_bytes(i64 n, int K, i64 m);
int backtransform(const TridiagWork& t, const double* Z, size_t ldZ, size_t sZ, i64 m, z2* C,
                  size_t ldC, size_t sC, void* work, hipStream_t s, float2* C32 = nullptr,
                  int* width_used = nullptr, int fixed_width = 0);

// gemm.hip: matrix-core GEMMs (C = alpha op(A) op(B) + beta C, batched with
// strides; Op::H = conjugate transpose, plain transpose for real data)
enum class Op { N, H };
int zgemm(Op ta, Op tb, i64 M, i64 N, i64 K, double alpha, const z2* A, i64 lda, size_t sA,
          const z2* B, i64 ldb, size_t sB, double beta, z2* C, i64 ldc, size_t sC, int batch,
          hipStream_t s);
int cgemm(Op ta, Op tb, i64 M, i64 N, i64 K, double alpha, const float2* A, i64 lda, size_t sA,
          const float2* B, i64 ldb, size_t sB, double beta, float2* C, i64 ldc, size_t sC,
          int batch, hipStream_t s);
int dgemm(Op ta, Op tb, i64 M, i64 N, i64 K, double alpha, const double* A, i64 lda, size_t sA,
          const double* B, i64 ldb, size_t sB, double beta, double* C, i64 ldc, size_t sC,
          int batch, hipStream_t s);
// R = A A^H (n x n, rank k) on the lower 128 x 128 tiles (diagonal tiles whole)
int herk_lower(i64 n, i64 k, const z2* A, i64 lda, z2* R, i64 ldr, hipStream_t s);
// rho_c(i,j) (+)= ph_c conj(R(i,j)) for all i, j and the nnc cells (stride n^2),
// R = A A^H; first: overwrite instead of accumulate
int herk_cells(i64 n, i64 k, const z2* A, i64 lda, int nnc, const double* ph, z2* rho,
               bool first, hipStream_t s);

// density.hip
int bulk_density(i64 n, int K, const z2* C, size_t ldC, size_t sC, i64 nocc, int nnc,
                 const double* phase_cells, z2* rho, bool accumulate, hipStream_t s);
int window_density(i64 n, int nnc, i64 S, const z2* Psi, const double* wph, z2* rho,
                   double scale, hipStream_t s, const i64* Sdev = nullptr);

// scf_device.hip: the steps of one SCF iteration that the host loop does on
// the host (sort, Fermi level, occupations, OptimalStep, convergence
// measure), as kernels, so that a whole iteration is one launch sequence.
constexpr int OCC_MAX = 4096;  // window states (nb x Nk) the device sort holds
struct OccArgs {
  int N, Nk, nb, nnc, numk, occ_mode;
  i64 nF;
  double energy_tol, kT;
  const double* band;     // [N] window energies, flat band-major b * Nk + k
  const double* ph_cell;  // [Nk][nnc][2] cell phases of each k
  int* idx;               // [N] window column (b + nb k) of each weighted state
  double* wph;            // [N][nnc][2] occupation x cell phase
  i64* S;                 // number of weighted states
  double* occ;            // E_F, nOcc, nPart, DegFactor
};
int device_occupations(const OccArgs& a, hipStream_t s);
int band_gather(i64 n, i64 nb, i64 nL1, int K, i64 k0, i64 Nk, const double* evals, double* band,
                hipStream_t s);
int gather_cols_dev(i64 n, i64 Smax, const i64* S, const int* idx, const z2* Cwin, z2* Psi,
                    hipStream_t s);
struct StepArgs {
  int mix_mode;
  double beta, step_alt, alpha, alphaH, U;
  const double* e5;    // energy-pass sums (K_a, K_b, F_aa, F_ab, F_bb)
  const double* dots;  // k_dots [0..5], k_bias_dot [8..9]
  const double* occ;   // E_F, nOcc, nPart
  double* st;          // state: [0..3] energies, [4] E_tot, [5] step
  double* row;         // trajectory row (12)
};
int scf_step_update(const StepArgs& a, hipStream_t s);
// row[6] = max_i |((1-s) da_i + s db_i) - da_i|, s = st[5]
int density_convergence(i64 n, const double* da, const double* db, const double* st, double* row,
                        hipStream_t s);
int mix_dev(size_t tot, z2* a, const z2* b, const double* st, hipStream_t s);
int potential_dev(i64 n, const double* ya, const double* yb, double* da, const double* db,
                  const double* st, double alphaH, double U, double* pot, hipStream_t s);

// c3.hip: the C3-reduced k-set (nenforceC3 = 1).
struct C3Geom {
  i64 n, per;        // sites, sites per layer
  int nl;            // layers
  const int* perm;   // device [2][n]: p, then p o p (0-based)
  const z2* tab;     // device corner factors [set][nl][2]
};
struct C3Col {       // one image column: dst(:, dst) = src(p^r(i), src) * corner(set, i)
  i64 src, dst;
  int r, set;
};
struct C3Slot {      // one orbit representative of a batch
  const double *ph0, *ph1, *ph2;  // device cell phases [nnc][2] of k, its C3 and C3^2 partners
  int fset, uset;                 // corner factor sets of the C3 and C3^2 block images
  double w;                       // 1, or 1/3 at a C3-invariant momentum
};
int c3_columns(const C3Geom& g, i64 ncol, const C3Col* d, const z2* src, z2* dst, hipStream_t s);
int bulk_density_c3(const C3Geom& g, int K, const z2* C, size_t ldC, size_t sC, i64 nocc, int nnc,
                    const C3Slot* slot, z2* rho, bool accumulate, hipStream_t s);
i64 c3_pairs(const double* C, i64 n, i64 nl, i64* perm);

// valley.hip: the valley projection (nenforceValley).
constexpr int VALLEY_MAXL = 8;  // layers
struct ValleyGeom {
  i64 n, per;                    // sites, sites per layer
  int nl, nnc;                   // layers, stored cells
  double Rp[4], Rm[4];           // RotMatrix (row-major) and its transpose
  signed char rot[VALLEY_MAXL];  // RotateLayers
  CellShifts cells;
  const double *X, *Y;           // device site coordinates
};
int valley_phase(const ValleyGeom& g, double vkx, double vky, z2* V, hipStream_t s);
size_t valley_workspace_bytes(i64 n, i64 nb, int K);
int valley_transform(const ValleyGeom& g, const double* vk, int K, i64 nb, z2* C, z2* V,
                     void* work, hipStream_t s);

// sym.hip: nenforceT / C2 / C2T on the nnc cells of rho (sym bits 1, 2, 4;
// scratch: one n x n matrix; perm: device C2 partners, 0-based).
int symmetrize(i64 n, int nnc, z2* rho, z2* scratch, int sym, const int* perm, hipStream_t s);
i64 c2_pairs(const double* C, i64 n, i64 nl, i64* perm);

}  // namespace forge
