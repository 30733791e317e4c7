/*
 * forge_hf.h -- C ABI of the MI355X Hartree-Fock engine (libforge_hip.so).
 *
 * The entry points replace the hot-path subroutines of the reference FORGE
 * code; each declaration names the reference interface it stands in for.
 * Conventions (the reference's, so a Fortran host binds them with
 * ISO_C_BINDING and passes its arrays unchanged):
 *   - arrays are column-major (Fortran order) host arrays;
 *   - int64_t    = integer(dp)   (Setup.f90:6, dp = 8)
 *   - double     = real(dp)
 *   - forge_z    = complex(dp) = double _Complex (two doubles, re then im)
 *   - return 0 on success, > 0 an eigensolver failure count (LAPACK "info"
 *     meaning), < 0 an argument or HIP runtime error (forge_last_error()
 *     describes it).
 * Module-level constants of Setup.f90 are passed once in forge_params.
 * Calls on one context must come from one host thread at a time.
 */
#ifndef FORGE_HF_H
#define FORGE_HF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } forge_z;
typedef struct { float re, im; } forge_c; /* complex(4) */

/* Runtime form of the Setup.f90 parameters (Setup.f90:6-74). */
typedef struct {
  int64_t ndim, nlayers, numk, numI, numC, nscreen, fphase, nnc;
  int64_t nLower, nUpper, numb;
  double nfilling, U, alpha, alphaH, Delta, xi, tz, a0, r0, d0;
  double energy_tol, step_alt;
} forge_params;

/* SCF options beyond the reference (defaults reproduce the reference:
 * occ_mode 0, mix_mode 0, etol 0). */
typedef struct {
  int32_t occ_mode;   /* 0: T=0 rule with the reference's zdeg = (float(DegFactor), 8) weight
                         (HartreeFock.f90:529); 1: T=0 rule with real DegFactor;
                         2: Fermi-Dirac occupations at temperature kT */
  int32_t mix_mode;   /* 0: optimal damping (OptimalStep, HartreeFock.f90:1148-1227); 1: linear */
  double kT;          /* eV, occ_mode 2 */
  double beta;        /* linear mixing weight, mix_mode 1 */
  double etol;        /* > 0: forge_scf_run stops once |E_in - E_new| < etol (eV per cell) */
  int32_t precision;  /* 0: FP64 everywhere; 1: mixed -- the Householder hemv streams an FP32
                         copy of the trailing matrix (FP64 accumulation) and the
                         back-transformation runs in complex FP32 (columns normalised in
                         FP64) until |E_in - E_new| < switch_de, then FP64 to convergence */
  int32_t timing;     /* nonzero: per-stage HIP-event timings (forge_scf_timings) */
  double switch_de;   /* eV, precision 1 (0: 1e-4) */
  int32_t graph;      /* 1: the Householder reduction's launches (3 per column and one
                         trailing update per panel) are captured into a HIP graph on first
                         use and replayed while the batch shape is unchanged; not with timing.
                         2: the whole SCF iteration runs on the device (sort, Fermi level,
                         occupations, OptimalStep / linear mixing and the convergence measure
                         as kernels; one host synchronisation per iteration); 3: as 2, each
                         phase's iteration (mixed, FP64) captured into one HIP graph on its
                         first use and replayed.  2 and 3 apply to the plain loop (no C3 set,
                         valley projection or symmetry enforcement, every k-point in one
                         eigensolver batch, nb * numk^2 <= 4096) without timing; the first
                         iteration of each phase of a solve runs on the host path. */
  int32_t reduction;  /* reduction to tridiagonal form: 0 by order (two-stage from n = 1024 on),
                         1 one-stage (tridiag.hip: a Householder matrix-vector product of the
                         trailing matrix per column), 2 two-stage (twostage.hip: band of width
                         32 by matrix-core products per panel, then bulge chasing); the
                         environment variable FORGE_REDUCTION=1 or 2 overrides it */
} forge_scf_options;

typedef struct forge_ctx forge_ctx;

/* Last error message of the calling thread (static storage). */
const char* forge_last_error(void);

/* Number of visible HIP devices (0 when none). */
int forge_device_count(void);

/* Uploads the geometry (Coords n x 3, tn 2 x 3 = [t1 t2 t3], the k grid of
 * SampleBZ, Geometry.f90:96-127, and the stored cells of OrderNeighborCells,
 * Geometry.f90:40-91) and builds the k-independent pair tables fv/ft of
 * every stored cell on the device.  The context owns all device memory. */
int forge_ctx_create(const forge_params* p, const double* Coords, const double* tn,
                     const double* MomentaValues, const int64_t* nMomentaComponents,
                     const int64_t* nUnitCell_1, const int64_t* nUnitCell_2, forge_ctx** out);
int forge_ctx_destroy(forge_ctx* ctx);

/* LongRangeInteraction(LongRange, Coords, ndim, t1, t2), HartreeFock.f90:14-53. */
int forge_longrange(forge_ctx* ctx, double* LongRange);

/* TightBindingHamiltonian(zH, Delta, Coords, ..., vk, tn), TightBinding.f90:13-55.
 * Lower triangle of zH (ndim x ndim) is set, the strict upper triangle is 0. */
int forge_tb_hamiltonian(forge_ctx* ctx, forge_z* zH, const double* vk);

/* HamiltonianHartreeFock(zH, Coords, Potential, alpha, Delta, zFock, ..., vk, tn),
 * HartreeFock.f90:57-127.  zFock is ndim x ndim x nnc.  Lower triangle only. */
int forge_hamiltonian_hf(forge_ctx* ctx, forge_z* zH, const double* Potential,
                         const forge_z* zFock, const double* vk);

/* zdiagRed(A, eigenvalues, 'V', il, iu), LapackRoutines.f90:381-433: the
 * eigenpairs il..iu (1-based, ascending) of the Hermitian matrix whose lower
 * triangle A (n x n) holds.  On return A(:, 1:iu-il+1) are the eigenvectors
 * and eigenvalues(1:iu-il+1) the eigenvalues.  *info as zheevr's. */
int forge_zdiagred(forge_z* A, int64_t n, double* eigenvalues, int64_t il, int64_t iu,
                   int32_t* info);

/* Products of the reference's LAPACK shim on the matrix-core GEMMs (host
 * arrays in and out; no context).
 * z_vector_mul(A, C, z_factor), LapackRoutines.f90:168-183: C (n x n) +=
 *   z A A^H with A n x k; the reference passes k = 1 (one rank-1 zgemm per
 *   band), k > 1 is the same update for k bands at once.  z_factor may be
 *   null (1).
 * z_matrix_mul(A, B, C), LapackRoutines.f90:208-220: C (m x n) = A (m x k) B.
 * z_matrix_overlap(A, B, C, z_factor), LapackRoutines.f90:119-139:
 *   C (m0 x m0) = z A^H B with A, B m1 x m0; z_factor may be null (1).
 * c_matrix_mul(A, B, C), LapackRoutines.f90:224-236: complex(4) C = A B. */
int forge_z_vector_mul(forge_z* C, const forge_z* A, int64_t n, int64_t k, const forge_z* z_factor);
int forge_z_matrix_mul(forge_z* C, const forge_z* A, const forge_z* B, int64_t m, int64_t k,
                       int64_t n);
int forge_z_matrix_overlap(forge_z* C, const forge_z* A, const forge_z* B, int64_t m1, int64_t m0,
                           const forge_z* z_factor);
int forge_c_matrix_mul(forge_c* C, const forge_c* A, const forge_c* B, int64_t m, int64_t k,
                       int64_t n);

/* No counterpart in the reference: forge_zdiagred for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  Same arguments and results; every
 * device buffer comes from a preallocated 32 MiB device array per GPU, so
 * n <= 512 (larger n returns -2).  Runs on the null stream.  Calls on one
 * GPU share that array: host threads are serialised by a lock; overlapping
 * calls by other means race. */
int forge_zdiagred_noalloc(forge_z* A, int64_t n, double* eigenvalues, int64_t il, int64_t iu,
                           int32_t* info);

/* Solve(zFockBulk, zFock, Potential, alpha, Bands, zEigenvectors, ...),
 * HartreeFock.f90:402-483: every k of the grid is assembled, diagonalised and
 * its bulk density accumulated; zFockBulk (ndim x ndim x nnc) is the
 * unnormalised sum over k, Bands (numb x Nk) and zEigenvectors
 * (ndim x numb x Nk) hold the window nLower..nUpper. */
int forge_solve(forge_ctx* ctx, forge_z* zFockBulk, const forge_z* zFock,
                const double* Potential, double* Bands, forge_z* zEigenvectors);

/* GetFock(zFock, zSortedEigenvectors, ndim, numb, numk, DegFactor, OccStates,
 * PartOccStates, ..., nSortedMomenta), HartreeFock.f90:487-560.  Adds the
 * window density to zFock (not cleared), with the reference's weights
 * (occ_mode of the last forge_scf_set_options; 0 by default). */
int forge_getfock(forge_ctx* ctx, forge_z* zFock, const forge_z* zSortedEigenvectors,
                  const int64_t* nSortedMomenta, double DegFactor, int64_t OccStates,
                  int64_t PartOccStates);

/* GetKineticEnergy / GetFockEnergy / GetHartreeHubbardEnergy
 * (HartreeFock.f90:1012-1145) in one pass: out = {Ekin, EF, EH, EU} for the
 * density matrix zFock and site densities dens (Density - DensitySub). */
int forge_energies(forge_ctx* ctx, const forge_z* zFock, const double* dens, double* out4);

/* ---- C3-reduced k-set (nenforceC3 = 1, Setup.f90:27) ----
 * RotateLayers: nlayers signs (Setup.f90:8); nC3pairs: ndim site indices,
 * 1-based, as C3_RelatedPoints returns them.  numk must be a multiple of 3
 * (the reference's invariant-momentum test names Gamma, K and K' by their
 * flat indices 1, numk/3 (numk+1) + 1, 2 numk/3 (numk+1) + 1). */

/* C3_RelatedPoints(nC3pairs, Coords, ndim), Geometry.f90:1172-1223, on the
 * host (no device work): nC3pairs(i) = the site of the same layer and
 * sublattice half within 1e-3 of the +120 degree image of site i; the two
 * corner sites of each layer map to themselves.  Returns the number of sites
 * matched (ndim for a C3-symmetric cell; unmatched entries are -1). */
int64_t forge_c3_pairs(const double* Coords, int64_t ndim, int64_t nlayers, int64_t* nC3pairs);

/* Solve_C3(zFockBulk, zFock, Potential, alpha, Bands, zEigenvectors, ...,
 * RotateLayers, ..., nC3pairs), HartreeFock.f90:131-398: H(k) is assembled
 * and diagonalised only at the first momentum of each C3 orbit; the partners
 * receive its window bands and permuted, corner-phased eigenvectors, and the
 * bulk density adds the orbit's block and its two C3 images (a third of each
 * at Gamma, K, K').  Outputs as forge_solve, for all Nk momenta. */
int forge_solve_c3(forge_ctx* ctx, forge_z* zFockBulk, const forge_z* zFock,
                   const double* Potential, double* Bands, forge_z* zEigenvectors,
                   const int64_t* RotateLayers, const int64_t* nC3pairs);

/* GetFock_C3(zFock, zSortedEigenvectors, ndim, RotateLayers, numb, numk,
 * DegFactor, OccStates, PartOccStates, ..., nSortedMomenta, nC3pairs),
 * HartreeFock.f90:564-881: as forge_getfock, except that a state at a
 * C3-invariant momentum adds (T + C3[T] + C3^2[T]) / 3 of its density T. */
int forge_getfock_c3(forge_ctx* ctx, forge_z* zFock, const forge_z* zSortedEigenvectors,
                     const int64_t* nSortedMomenta, double DegFactor, int64_t OccStates,
                     int64_t PartOccStates, const int64_t* RotateLayers,
                     const int64_t* nC3pairs);

/* ---- Symmetry enforcement (nenforceT / nenforceC2 / nenforceC2T,
 * Setup.f90:28-30) ----
 * flags: 1 = T, 2 = C2, 4 = C2T (any combination, applied in that order);
 * nC2pairs: ndim site indices, 1-based, as C2_RelatedPoints returns them
 * (needed for C2 and C2T). */

/* C2_RelatedPoints(nC2pairs, Coords, ndim), Geometry.f90:1140-1167, on the
 * host (no device work): within each layer, sites of the two sublattice
 * halves at r and -r (within 1e-5) are partners.  Returns the number of pairs
 * (ndim/2 for a C2-symmetric cell; unmatched entries are -1). */
int64_t forge_c2_pairs(const double* Coords, int64_t ndim, int64_t nlayers, int64_t* nC2pairs);

/* The symmetrisations of Main.f90:514-534 on a density (ndim x ndim x nnc,
 * in place, on the device): T keeps the real part; C2 sets
 * zFock = (zFock + conjg(transpose(zFock(nC2pairs, nC2pairs))))/2 and C2T the
 * same without conjg, on every cell. */
int forge_symmetrize(forge_ctx* ctx, forge_z* zFock, int32_t flags, const int64_t* nC2pairs);

/* ---- Valley projection (nenforceValley, Setup.f90:31) ----
 * RotateLayers as for the C3 entry points; RotMatrix (Main.f90:133) follows
 * from the cell (aMoire = ndim / (2 nlayers)). */

/* ValleyPhase(zV, Coords, nUnitCell_1, nUnitCell_2, ndim, numNeighborCells,
 * vk, tn, RotMatrix), TightBinding.f90:60-165 (fphase = 0): the valley
 * operator at momentum vk (ndim x ndim). */
int forge_valley_phase(forge_ctx* ctx, const double* vk, const int64_t* RotateLayers, forge_z* zV);

/* ValleyTransform(zEigenvectors, ...), TightBinding.f90:170-211: per k,
 * zEigenvectors(:, :, k) (ndim x numb, all Nk in place) <- C sign(C^H V C)
 * (unitarize, LapackRoutines.f90:865-893). */
int forge_valley_transform(forge_ctx* ctx, forge_z* zEigenvectors, const int64_t* RotateLayers);

/* GetHartreeHubbardEnergy(fsumH, fsumU, LongRange, dens1, dens2, ndim),
 * HartreeFock.f90:1066-1092, with the context's device copy of LongRange:
 * out2 = {1/2 (d1+d2).L.(d1+d2), d1.d2}. */
int forge_hartree_hubbard(forge_ctx* ctx, const double* dens1, const double* dens2, double* out2);

/* ---- device-resident SCF (the loop of Main.f90:364-679 on the device) ---- */

int forge_scf_set_options(forge_ctx* ctx, const forge_scf_options* opt);

/* nenforceC3 = 1 for the device-resident loop: forge_scf_init and
 * forge_scf_run call Solve_C3 / GetFock_C3 where the reference does
 * (Main.f90:244-250, 273-276, 375-378, 500-505).  Both pointers null: back to
 * Solve / GetFock.  Not available in sweeps. */
int forge_scf_set_c3(forge_ctx* ctx, const int64_t* RotateLayers, const int64_t* nC3pairs);

/* The symmetrisations of forge_symmetrize applied by forge_scf_run to every
 * iteration's new density, where Main.f90:514-534 applies them (flags 0: off). */
int forge_scf_set_symmetry(forge_ctx* ctx, int32_t flags, const int64_t* nC2pairs);

/* nenforceValley = 1 for the device-resident loop (Main.f90:469-494): each
 * iteration's window density is the mean of those of the states and of their
 * ValleyTransform.  Null: off.  Not with the C3-reduced k-set. */
int forge_scf_set_valley(forge_ctx* ctx, const int64_t* RotateLayers);

/* Initial state: rho_start (ndim x ndim x nnc) as read from a checkpoint
 * (Main.f90:187-235), or NULL for the first diagonalisation
 * (Main.f90:238-288).  Computes densities, energies and the potential
 * (Main.f90:290-346). */
int forge_scf_init(forge_ctx* ctx, const forge_z* rho_start);

/* Initial state from a proposed density (the learned guess; no reference
 * counterpart, it enters at the initial-state block Main.f90:187-288): one
 * diagonalisation of H[rho_guess] at every k, and the state is the occupied
 * density it yields, built as the first diagonalisation builds it from the
 * tight-binding Hamiltonian (Main.f90:238-282); then as forge_scf_init. */
int forge_scf_init_guess(forge_ctx* ctx, const forge_z* rho_guess);

/* The same starting from the current state's density on the device (normally
 * the first diagonalisation of forge_scf_init(ctx, NULL)): its cell-0
 * diagonal is set to diag (ndim values) and the nupd distinct entries
 * (row, col, cell), 0-based, are incremented by delta; no ndim^2 host copy. */
int forge_scf_init_guess_update(forge_ctx* ctx, const double* diag, int64_t nupd,
                                const int64_t* row, const int64_t* col, const int64_t* cell,
                                const forge_z* delta);

/* As forge_scf_init_guess_update, but the updated density itself becomes the
 * start state (its potential and energies are formed; no diagonalisation):
 * the first forge_scf_run iteration solves H[guess] and mixes it with the
 * guess.  No reference counterpart (learned initial guess). */
int forge_scf_init_update(forge_ctx* ctx, const double* diag, int64_t nupd, const int64_t* row,
                          const int64_t* col, const int64_t* cell, const forge_z* delta);

/* Runs up to max_iter SCF iterations (Main.f90:365-679).  traj receives one
 * row of 12 doubles per iteration: it, E_in, 2Ekin_in, 2EF_in, EH_in, EU_in,
 * max|dDensity|, Step, E_in - E_new, E_Fermi, nOcc, nPart.  *iters_done is
 * set.  Stops early once |E_in - E_new| < opt.etol when etol > 0. */
int forge_scf_run(forge_ctx* ctx, int64_t max_iter, double* traj, int64_t* iters_done);

/* Copies the current density matrix (ndim x ndim x nnc) to the host. */
int forge_scf_get_density(forge_ctx* ctx, forge_z* rho);

/* ---- Two spins (numS = 2) on the device ----
 * The SCF loop of Main.f90:236-679 with numS = 2 (the per-spin Solve and
 * SortEnergies_1Spin, SortEnergies_2Spins and the two-spin Fermi block of
 * Main.f90:416-458, GetFock per spin, OptimalStep on the two-spin energy of
 * HartreeFock.f90:1148-1227, the potentials of Main.f90:586-595), with both
 * spins' densities, states and potentials kept on the device.  Replaces the
 * per-spin host-array calls of the drop-in host's numS = 2 path.  spin1 and
 * spin2 are two contexts of the same system (same parameters and
 * coordinates), owned by the caller; each carries its spin's density, bulk
 * and window states and eigensolver batch.  Their forge_scf_set_options must
 * agree, with occ_mode 0 or 1 (the reference's T = 0 rule), mix_mode 0
 * (OptimalStep), graph 0 or 1 and no timing; forge_scf_set_c3 and
 * forge_scf_set_symmetry, if used, on both (the valley projection is
 * refused: it runs with one spin). */
typedef struct forge_scf2 forge_scf2;
int forge_scf2_create(forge_ctx* spin1, forge_ctx* spin2, forge_scf2** out);
int forge_scf2_destroy(forge_scf2* pair);
/* rho_start: ndim x ndim x nnc x 2 (both spins, the reference's two
 * checkpoint files, Main.f90:201-235), or NULL: the first diagonalisation of
 * spin 1 with zero density and potential, copied to spin 2 (Main.f90:238-288).
 * Returns eigensolver failures (> 0) or an error (< 0). */
int forge_scf2_init(forge_scf2* pair, const forge_z* rho_start);
/* Up to max_iter iterations.  traj rows (12 doubles): it, E_in, Ekin1+Ekin2,
 * EF1+EF2, EH, EU (input state; the reference's numS = 2 Energy file,
 * Main.f90:632-638), max over spins of max|dDensity|, Step, E_in - E_new,
 * E_Fermi, nOcc1+nOcc2, nPart1+nPart2.  Stops once |E_in - E_new| < etol
 * (etol > 0); with precision 1 the mixed phase runs until switch_de. */
int forge_scf2_run(forge_scf2* pair, int64_t max_iter, double* traj, int64_t* iters_done);
/* Both spins' densities, ndim x ndim x nnc x 2. */
int forge_scf2_get_density(forge_scf2* pair, forge_z* rho);

/* ---- Twist-angle sweeps (no reference counterpart: the reference runs one
 * system per process, Main.f90:34-685) ----
 * A sweep advances the SCF of nsys systems of different orders together (e.g.
 * TBG cells m = 8..14, each with its own filling): per iteration all their
 * H(k) are reduced to tridiagonal form in ONE batch (uneven sizes; see
 * TridiagWork::first_col), then every system finishes its iteration as
 * forge_scf_run would.  The contexts stay owned by the caller; each needs
 * forge_scf_set_options and forge_scf_init before forge_sweep_run, and must
 * hold all its k-points in one eigensolver batch.  With precision 1 the shared
 * reduction streams an FP32 copy while every system in the batch is in its
 * mixed phase; each system's back-transformation and switch to FP64 follow
 * its own options. */
typedef struct forge_sweep forge_sweep;
int forge_sweep_create(forge_ctx* const* ctxs, int64_t nsys, forge_sweep** out);
int forge_sweep_destroy(forge_sweep* sweep);

/* Up to max_iter iterations of every system; a system stops once it
 * converges (its opt.etol) and leaves the batch.  traj (nsys x max_iter x 12,
 * may be NULL) gets forge_scf_run's rows per system, iters_done (nsys) the
 * iterations each system ran. */
int forge_sweep_run(forge_sweep* sweep, int64_t max_iter, double* traj, int64_t* iters_done);

/* Per-stage device time of the last forge_scf_run (milliseconds, HIP events):
 * [0] assembly, [1] tridiagonal reduction, [2] tridiagonal eigensolver,
 * [3] back-transformation, [4] density products, [5] energies + mixing,
 * [6] total, [7] launches of the assembly kernel, [8] summed time of the
 * assembly kernel (ms), [9] timed launches of the Householder hemv kernel
 * (every 17th column), [10] their summed time (ms), [11] their algorithmic
 * bytes (8 m^2 per matrix in FP64, 4 m^2 from the FP32 copy; m the trailing
 * order), [12] iterations run with precision 1 still in its mixed phase,
 * [13] block width of the last back-transformation (tuned per shape),
 * [14] the Householder hemv's variants of the last reduction (tuned per
 * shape), one decimal digit per band of trailing orders (thousands
 * m >= 6144, hundreds 3072..6143, tens 1536..3071, units below 1536): 1 and
 * 2 segments of 1024 and 2048 rows with 16 lanes per column chunk, 3 and 4
 * the same with 8 lanes and 16-byte row-pair loads (FP32 copy only).
 * FORGE_HEMV_FORCE="<segment>,<layout>" in the environment forces one
 * variant (layout 0 or 1) in every band.  With the two-stage reduction:
 * [15] timed launches of the chase's back-transformation kernel (every 4th
 * group of sweeps), [16] their summed time (ms), [17] their algorithmic flops
 * (16 per reflector element per eigenvector column), [18] the bytes of the
 * eigenvector rows they pass over, read and written once.  [19..23] are
 * zero. */
#define FORGE_NTIMINGS 24
int forge_scf_timings(forge_ctx* ctx, double* out24);

/* ---- MLP of the learned initial guess (no reference counterpart) ----
 * 3 layers of width 128 with ReLU: Y = relu(relu(X W1 + b1) W2 + b2) W3 + b3,
 * X (N x din, row-major, din <= 32), Y (N x dout, dout <= 16), parameters
 * packed [W1 (din x 128), b1, W2 (128 x 128), b2, W3 (128 x dout), b3], W
 * row-major (in x out), FP32. */
int64_t forge_mlp_param_count(int64_t din, int64_t dout);

/* Masked mean-squared loss sum(mask (Y - T)^2) / sum(mask) of the N rows and
 * its gradient with respect to the parameters (FP32 kernels). */
int forge_mlp_loss_grad(const float* params, int64_t din, int64_t dout, const float* X,
                        const float* T, const float* mask, int64_t N, float* loss, float* grad);

/* Adam (lr, beta1 0.9, beta2 0.999, eps 1e-8) on mini-batches of `batch` rows
 * drawn without replacement in epochs (order from splitmix64(seed)); params
 * in/out; loss_hist (steps, may be NULL) gets each step's batch loss. */
int forge_mlp_train(float* params, int64_t din, int64_t dout, const float* X, const float* T,
                    const float* mask, int64_t N, int64_t steps, int64_t batch, double lr,
                    uint64_t seed, float* loss_hist);

/* Y = network(X): bf16 != 0 on the matrix cores with bf16 operands and FP32
 * accumulation, else the FP32 training kernels. */
int forge_mlp_infer(const float* params, int64_t din, int64_t dout, const float* X, int64_t N,
                    int32_t bf16, float* Y);

#ifdef __cplusplus
}
#endif
#endif
