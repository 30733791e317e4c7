/*
 * forge_oracle.c -- CPU restatement of the reference FORGE Hartree-Fock hot
 * path.  TEST INFRASTRUCTURE ONLY: it is the checker for the HIP library and
 * the timed CPU baseline ("port"); nothing in the product links or calls it.
 *
 * Every function follows one reference routine, cited file:line, with the
 * same loop order, the same LAPACK driver (zheevr range 'I', uplo 'L',
 * abstol = dlamch('S')), the same rank-1 zgemm bulk-density loop and OpenMP
 * over k-points.  It is pinned by tests/test_oracle.py against the fixtures
 * that the compiled reference wrote (the golden .npz and .traj.txt files) and by
 * closed forms (ft/fv special values, Hermiticity, projector identities).
 *
 * Layout: every array is column-major (Fortran order) exactly as the
 * reference declares it; complex values are C99 double complex
 * (= Fortran complex(8)); indices are int64 (= integer(dp), Setup.f90:6).
 *
 * The reference's variants are pinned the same way, against the reference
 * program built with the setting (trajectory files under tests/golden/):
 * the C3-reduced k-set (nenforceC3 = 1: orc_c3_pairs, orc_solve_c3,
 * orc_getfock_c3; <case>_c3), two spins (numS = 2: orc_sort_2spins,
 * orc_fermi_2spins, orc_scf2; <case>_s2), the T / C2 / C2T symmetrisations
 * (orc_c2_pairs, orc_symmetrize) and combinations of them.
 *
 * Extensions with no reference counterpart (used by the configurations that
 * ask for them): Fermi-Dirac window occupations (occ_mode 2), physical
 * degenerate-shell weight (occ_mode 1), linear mixing (mix_mode 1) and an
 * energy stopping tolerance.  occ_mode 0 + mix_mode 0 + etol 0 is the
 * reference behaviour.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef int64_t i64;
typedef double complex zc;

/* Runtime form of the Setup.f90 parameters (Setup.f90:6-74). */
typedef struct {
  i64 ndim, nlayers, numk, numI, numC, nscreen, fphase, nnc;
  i64 nLower, nUpper, numb;
  double nfilling, U, alpha, alphaH, Delta, xi, tz, a0, r0, d0;
  double energy_tol, step_alt;
} orc_params;

extern void zheevr_(const char*, const char*, const char*, const int*, zc*, const int*,
                    const double*, const double*, const int*, const int*, const double*,
                    int*, double*, zc*, const int*, int*, zc*, const int*, double*,
                    const int*, int*, const int*, int*, size_t, size_t, size_t);
extern void zgemm_(const char*, const char*, const int*, const int*, const int*, const zc*,
                   const zc*, const int*, const zc*, const int*, const zc*, zc*, const int*,
                   size_t, size_t);
extern double dlamch_(const char*, size_t);
extern void zheev_(const char*, const char*, const int*, zc*, const int*, double*, zc*, const int*,
                   double*, int*, size_t, size_t);

static const double PI = 3.14159265358979323846264338327950288;

/* ------------------------------------------------------------------ pair functions */

/* ft: Slater-Koster pz hopping, TightBinding.f90:216-230. */
double orc_ft(const orc_params* p, const double* ri, const double* rj) {
  double dx = ri[0] - rj[0], dy = ri[1] - rj[1], dz = ri[2] - rj[2];
  double r = sqrt(dx * dx + dy * dy + dz * dz);
  if (!(r > 0.001)) return 0.0;
  double cs = dz / r;
  return -2.7 * exp((p->a0 - r) / p->r0) * (1.0 - cs * cs) + 0.48 * exp((p->d0 - r) / p->r0) * cs * cs;
}

/* fv: gate-screened Coulomb, HartreeFock.f90:1232-1274. */
double orc_fv(const orc_params* p, const double* ri, const double* rj) {
  double xi_ = ri[0], yi = ri[1], zi = ri[2], xj = rj[0], yj = rj[1], zj = rj[2];
  double dx = xi_ - xj, dy = yi - yj, dz = zi - zj;
  double r = sqrt(dx * dx + dy * dy + dz * dz);
  double r2 = sqrt(dx * dx + dy * dy);
  double fv = 0.0;
  if (p->nscreen == 2) {
    if (r > 0.01) {
      if (r2 < p->xi || r2 < 2.0 * (double)p->nlayers * p->tz) {
        for (int n = -20; n <= 20; ++n) {
          double sgn = (n & 1) ? -1.0 : 1.0;
          double zz = zj - ((double)n * p->xi + zi * sgn);
          fv += sgn / sqrt((xj - xi_) * (xj - xi_) + (yj - yi) * (yj - yi) + zz * zz);
        }
      } else {
        fv = 2.0 * sqrt(2.0) / sqrt(r2 / p->xi) * exp(-PI * r2 / p->xi) / p->xi;
      }
    }
  }
  if (p->nscreen == 1) {
    if (r > 0.01) {
      double zz = zi + zj + p->xi;
      fv = 1.0 / r - 1.0 / sqrt(dx * dx + dy * dy + zz * zz);
    } else {
      fv = 0.0;
    }
  }
  return fv;
}

/* ------------------------------------------------------------------ geometry */

/* NumberNeighborCells, Geometry.f90:14-35. */
i64 orc_num_neighbor_cells(i64 numI) {
  i64 c = 0;
  for (i64 n1 = -numI; n1 <= numI; ++n1)
    for (i64 n2 = -numI; n2 <= numI; ++n2)
      if (n1 + n2 <= numI && n1 + n2 >= -numI && (n1 != 0 || n2 != 0)) ++c;
  return c / 2 + 1;
}

/* OrderNeighborCells, Geometry.f90:40-91. */
void orc_order_neighbor_cells(i64 numI, i64 nnc, i64* u1, i64* u2) {
  for (i64 i = 0; i < nnc; ++i) u1[i] = u2[i] = 0;
  i64 ic = 0; /* 0-based position of the last filled cell */
  if (numI > 0) {
    u1[1] = -1; u2[1] = 0;
    u1[2] = -1; u2[2] = 1;
    u1[3] = 0;  u2[3] = -1;
    ic = 3;
  }
  if (numI > 1) {
    for (i64 n1 = -numI; n1 <= numI; ++n1)
      for (i64 n2 = -numI; n2 <= numI; ++n2)
        if (n1 + n2 <= numI && n1 + n2 >= -numI && ic + 1 < nnc && (n1 != 0 || n2 != 0)) {
          ++ic;
          u1[ic] = n1; u2[ic] = n2;
          if ((n1 == -1 && n2 == 0) || (n1 == -1 && n2 == 1) || (n1 == 0 && n2 == -1)) --ic;
        }
  }
}

i64 orc_ndim(i64 ntheta, i64 nlayers) {
  return nlayers * 2 * (ntheta * ntheta + (ntheta + 1) * (ntheta + 1) + (ntheta + 1) * ntheta);
}

/* WignerSeitzCell, Geometry.f90:132-273 (unrotated t1, t2). */
static void wigner_seitz(double* C, i64 ndim, i64 ntheta, i64 nlayers, const i64* rot,
                         const double* t1, const double* t2, double cs, double sn, double tz) {
  const double a1[2] = {0.5, sqrt(3.0) * 0.5}, a2[2] = {-0.5, sqrt(3.0) * 0.5};
  const double sq3 = sqrt(3.0);
  const double eps6 = (double)1e-6f; /* default-real literal in the reference */
  i64 nrad = 3 * ntheta;
  double rMax = 3.0 * (double)(ntheta * ntheta) + 3.0 * (double)ntheta + 1.0;
  double A = 3.0 * (double)ntheta + 1, B = 3.0 * (double)ntheta + 2;
  i64 nind = 0;
#define PUT(x, y, z) do { C[nind] = (x); C[nind + ndim] = (y); C[nind + 2 * ndim] = (z); ++nind; } while (0)
  for (i64 l = 1; l <= nlayers; ++l) {
    double z = (double)((float)(l - 1) - (float)(nlayers - 1) / 2.0f) * tz;
    if (rot[l - 1] == -1) {
      for (i64 n1 = -nrad; n1 <= nrad; ++n1)
        for (i64 n2 = -nrad; n2 <= nrad; ++n2) {
          double an1 = (n1 + 1.0 / 3.0), an2 = (n2 - 2.0 / 3.0);
          double r1 = fabs(an1 * A + an2 * B) + eps6, r2 = fabs(an1 - an2 * A) + eps6, r3 = fabs(an1 * B + an2) + eps6;
          if (r1 < rMax && r2 < rMax && r3 < rMax)
            PUT((n1 + 1.0 / 3.0) * a1[0] + (n2 - 2.0 / 3.0) * a2[0], (n1 + 1.0 / 3.0) * a1[1] + (n2 - 2.0 / 3.0) * a2[1], z);
        }
      PUT((-t2[0] - t1[0]) / 3.0, (-t2[1] - t1[1]) / 3.0, z);
      for (i64 n1 = -nrad; n1 <= nrad; ++n1)
        for (i64 n2 = -nrad; n2 <= nrad; ++n2) {
          double bn1 = n1 + 2.0 / 3.0, bn2 = n2 - 1.0 / 3.0;
          double r1 = fabs(bn1 * A + bn2 * B) + eps6, r2 = fabs(bn1 - bn2 * A) + eps6, r3 = fabs(bn1 * B + bn2) + eps6;
          if (r1 < rMax && r2 < rMax && r3 < rMax)
            PUT(bn1 * a1[0] + bn2 * a2[0], bn1 * a1[1] + bn2 * a2[1], z);
        }
      PUT((t2[0] + t1[0]) / 3.0, (t2[1] + t1[1]) / 3.0, z);
    } else if (rot[l - 1] == 1) {
      for (i64 n1 = -nrad; n1 <= nrad; ++n1)
        for (i64 n2 = -nrad; n2 <= nrad; ++n2) {
          double an1 = (n1 + 1.0 / 3.0) * (cs - sn / sq3) - 2 * (n2 - 2.0 / 3.0) * sn / sq3;
          double an2 = (n2 - 2.0 / 3.0) * (cs + sn / sq3) + 2 * (n1 + 1.0 / 3.0) * sn / sq3;
          double r1 = fabs(an1 * A + an2 * B) + eps6, r2 = fabs(an1 - an2 * A) + eps6, r3 = fabs(an1 * B + an2) + eps6;
          if (r1 < rMax && r2 < rMax && r3 < rMax)
            PUT(an1 * a1[0] + an2 * a2[0], an1 * a1[1] + an2 * a2[1], z);
        }
      PUT((t1[0] + t2[0]) / 3.0, (t1[1] + t2[1]) / 3.0, z);
      for (i64 n1 = -nrad; n1 <= nrad; ++n1)
        for (i64 n2 = -nrad; n2 <= nrad; ++n2) {
          double bn1 = (n1 + 2.0 / 3.0) * (cs - sn / sq3) - 2 * (n2 - 1.0 / 3.0) * sn / sq3;
          double bn2 = (n2 - 1.0 / 3.0) * (cs + sn / sq3) + 2 * (n1 + 2.0 / 3.0) * sn / sq3;
          double r1 = fabs(bn1 * A + bn2 * B) + eps6, r2 = fabs(bn1 - bn2 * A) + eps6, r3 = fabs(bn1 * B + bn2) + eps6;
          if (r1 < rMax && r2 < rMax && r3 < rMax)
            PUT(bn1 * a1[0] + bn2 * a2[0], bn1 * a1[1] + bn2 * a2[1], z);
        }
      PUT((-t1[0] - t2[0]) / 3.0, (-t1[1] - t2[1]) / 3.0, z);
    }
  }
#undef PUT
  if (nind != ndim) fprintf(stderr, "orc: WignerSeitzCell placed %lld sites, expected %lld\n", (long long)nind, (long long)ndim);
}

/* Geometry block of Main.f90:123-154 plus SampleBZ (Geometry.f90:96-127). */
int orc_geometry(i64 ntheta, i64 nlayers, const i64* rot, i64 numk, i64 numI, double tz,
                 double* Coords, double* tn, double* g1o, double* g12o, double* MomVal,
                 i64* nMomComp, i64* u1, i64* u2) {
  const double a1[2] = {0.5, sqrt(3.0) * 0.5}, a2[2] = {-0.5, sqrt(3.0) * 0.5};
  i64 ndim = orc_ndim(ntheta, nlayers);
  i64 nnc = orc_num_neighbor_cells(numI);
  orc_order_neighbor_cells(numI, nnc, u1, u2);
  double aM = 3.0 * (double)(ntheta * ntheta) + 3.0 * (double)ntheta + 1.0;
  double g1[2], g12[2];
  for (int c = 0; c < 2; ++c) {
    g1[c] = (4.0 * PI / 3.0) / aM * ((double)(3 * ntheta + 1) * a1[c] + a2[c]);
    g12[c] = (4.0 * PI / 3.0) / aM * ((double)(3 * ntheta + 2) * a2[c] - a1[c]);
  }
  double cs = 1.0 - 1.0 / (2.0 * aM), sn = sqrt(1.0 - cs * cs);
  double h = 0.5 * acos(cs), ch = cos(h), shh = sin(h);
  /* Rot = [[c, s], [-s, c]] (reshape([c,-s,s,c],[2,2]) column-major) */
#define ROT(v, o) do { double _x = (v)[0], _y = (v)[1]; (o)[0] = ch * _x + shh * _y; (o)[1] = -shh * _x + ch * _y; } while (0)
  ROT(g1, g1); ROT(g12, g12);
  for (i64 i1 = 0; i1 < numk; ++i1)
    for (i64 i2 = 0; i2 < numk; ++i2) {
      i64 ic = i1 * numk + i2;
      nMomComp[ic] = i1; nMomComp[ic + numk * numk] = i2;
    }
  for (i64 i2 = 0; i2 < numk; ++i2)
    for (i64 i1 = 0; i1 < numk; ++i1)
      for (int c = 0; c < 2; ++c)
        MomVal[i1 + numk * i2 + numk * numk * c] = ((double)i1 * g1[c] + (double)i2 * g12[c]) / (double)numk;
  double t1[2], t2[2];
  for (int c = 0; c < 2; ++c) {
    t1[c] = (double)ntheta * a1[c] + (double)(ntheta + 1) * a2[c];
    t2[c] = (double)(-ntheta - 1) * a1[c] + (double)(2 * ntheta + 1) * a2[c];
  }
  wigner_seitz(Coords, ndim, ntheta, nlayers, rot, t1, t2, cs, sn, tz);
  for (i64 n = 0; n < ndim; ++n) {
    double v[2] = {Coords[n], Coords[n + ndim]};
    ROT(v, v);
    Coords[n] = v[0]; Coords[n + ndim] = v[1];
  }
  ROT(t1, t1); ROT(t2, t2);
#undef ROT
  tn[0] = t1[0]; tn[1] = t1[1]; tn[2] = t2[0]; tn[3] = t2[1]; tn[4] = t2[0] - t1[0]; tn[5] = t2[1] - t1[1];
  g1o[0] = g1[0]; g1o[1] = g1[1]; g12o[0] = g12[0]; g12o[1] = g12[1];
  return 0;
}

static inline void site(const double* C, i64 n, i64 i, double* r) {
  r[0] = C[i]; r[1] = C[i + n]; r[2] = C[i + 2 * n];
}

/* ------------------------------------------------------------------ LongRange */

/* LongRangeInteraction, HartreeFock.f90:14-53. */
void orc_longrange(const orc_params* p, const double* C, const double* tn, double* L) {
  i64 n = p->ndim, nc = p->numC;
  const double* t1 = tn; const double* t2 = tn + 2;
#pragma omp parallel for schedule(dynamic, 4)
  for (i64 a = 0; a < n; ++a) {
    double ra[3], rb[3], rs[3];
    site(C, n, a, ra);
    for (i64 b = a; b < n; ++b) {
      site(C, n, b, rb);
      double s = 0.0;
      for (i64 n1 = -nc; n1 <= nc; ++n1)
        for (i64 n2 = -nc; n2 <= nc; ++n2)
          if (n1 + n2 <= nc && n1 + n2 >= -nc) {
            rs[0] = ra[0] - n1 * t1[0] - n2 * t2[0];
            rs[1] = ra[1] - n1 * t1[1] - n2 * t2[1];
            rs[2] = ra[2];
            s += orc_fv(p, rs, rb);
          }
      L[a + n * b] = s;
    }
  }
  for (i64 a = 0; a < n - 1; ++a)
    for (i64 b = a + 1; b < n; ++b) L[b + n * a] = L[a + n * b];
}

/* ------------------------------------------------------------------ Hamiltonian */

static inline double layer_bias(const orc_params* p, i64 i) {
  if (p->nlayers <= 1) return 0.0;
  i64 per = p->ndim / p->nlayers;
  i64 l = i / per;
  return (double)((float)l - (float)(p->nlayers - 1) / 2.0f) * p->Delta / (double)(float)(p->nlayers - 1);
}

/* HamiltonianHartreeFock, HartreeFock.f90:57-127; with zFock == NULL and
 * Potential == NULL it is TightBindingHamiltonian, TightBinding.f90:13-55.
 * Only the lower triangle (j <= i) is written; the rest is zeroed. */
void orc_hamiltonian_hf(const orc_params* p, zc* zH, const double* C, const double* Potential,
                        const zc* zFock, const i64* u1, const i64* u2, const double* vk,
                        const double* tn) {
  i64 n = p->ndim, nnc = p->nnc;
  const double* t1 = tn; const double* t2 = tn + 2;
  memset(zH, 0, sizeof(zc) * n * n);
#pragma omp parallel for schedule(dynamic, 8)
  for (i64 i = 0; i < n; ++i) {
    double ri[3], rj[3], rs[3];
    site(C, n, i, ri);
    if (Potential) zH[i + n * i] += Potential[i];
    for (i64 j = 0; j <= i; ++j) {
      site(C, n, j, rj);
      double ph = -(vk[0] * (ri[0] - rj[0]) + vk[1] * (ri[1] - rj[1])) * (double)p->fphase;
      zc zphase = cexp(I * ph);
      zc h = zH[i + n * j];
      if (zFock) h += -p->alpha * orc_fv(p, ri, rj) * conj(zFock[i + n * j]) * zphase;
      h += orc_ft(p, ri, rj) * zphase;
      for (i64 c = 1; c < nnc; ++c) {
        double Tx = u1[c] * t1[0] + u2[c] * t2[0], Ty = u1[c] * t1[1] + u2[c] * t2[1];
        double th = vk[0] * Tx + vk[1] * Ty;
        rs[0] = rj[0] - Tx; rs[1] = rj[1] - Ty; rs[2] = rj[2];
        h += orc_ft(p, ri, rs) * cexp(-I * th) * zphase;
        rs[0] = rj[0] + Tx; rs[1] = rj[1] + Ty;
        h += orc_ft(p, ri, rs) * cexp(I * th) * zphase;
      }
      if (zFock) {
        zc zs = 0;
        for (i64 c = 1; c < nnc; ++c) {
          double Tx = u1[c] * t1[0] + u2[c] * t2[0], Ty = u1[c] * t1[1] + u2[c] * t2[1];
          double th = vk[0] * Tx + vk[1] * Ty;
          rs[0] = rj[0] - Tx; rs[1] = rj[1] - Ty; rs[2] = rj[2];
          zs += -p->alpha * orc_fv(p, ri, rs) * conj(zFock[i + n * j + n * n * c]) * cexp(-I * th);
          rs[0] = ri[0] - Tx; rs[1] = ri[1] - Ty; rs[2] = ri[2];
          zs += -p->alpha * orc_fv(p, rj, rs) * zFock[j + n * i + n * n * c] * cexp(I * th);
        }
        h += zs * zphase;
      }
      zH[i + n * j] = h;
    }
  }
  for (i64 i = 0; i < n; ++i) zH[i + n * i] += layer_bias(p, i);
}

/* ------------------------------------------------------------------ eigensolver */

/* zdiagRed, LapackRoutines.f90:381-433: zheevr('V','I','L', il..iu),
 * abstol = dlamch('S'); A(:, 1:iu-il+1) <- eigenvectors. Returns info. */
int orc_zdiagred(zc* A, i64 n64, double* w, i64 il64, i64 iu64) {
  int n = (int)n64, il = (int)il64, iu = (int)iu64, m = 0, info = 0;
  int ncol = iu - il + 1;
  double vl = 0, vu = 0, abstol = dlamch_("S", 1);
  int* isuppz = (int*)malloc(sizeof(int) * 2 * (size_t)ncol);
  zc* z = (zc*)malloc(sizeof(zc) * (size_t)n * ncol);
  zc wq; double rq; int iq; int lw = -1, lrw = -1, liw = -1;
  zheevr_("V", "I", "L", &n, A, &n, &vl, &vu, &il, &iu, &abstol, &m, w, z, &n, isuppz,
          &wq, &lw, &rq, &lrw, &iq, &liw, &info, 1, 1, 1);
  lw = (int)creal(wq); lrw = (int)rq; liw = iq;
  zc* work = (zc*)malloc(sizeof(zc) * (size_t)lw);
  double* rwork = (double*)malloc(sizeof(double) * (size_t)lrw);
  int* iwork = (int*)malloc(sizeof(int) * (size_t)liw);
  zheevr_("V", "I", "L", &n, A, &n, &vl, &vu, &il, &iu, &abstol, &m, w, z, &n, isuppz,
          work, &lw, rwork, &lrw, iwork, &liw, &info, 1, 1, 1);
  memcpy(A, z, sizeof(zc) * (size_t)n * ncol);
  free(work); free(rwork); free(iwork); free(isuppz); free(z);
  return info;
}

/* z_vector_mul, LapackRoutines.f90:168-183: C <- C + zf * a a^H (zgemm k=1). */
static void vector_mul(const zc* a, zc* Cm, int n, zc zf) {
  const int one = 1; const zc beta = 1.0;
  zgemm_("N", "C", &n, &n, &one, &zf, a, &n, a, &n, &beta, Cm, &n, 1, 1);
}

static inline double cell_phase_arg(const orc_params* p, i64 ivk1, i64 ivk2, i64 n1, i64 n2) {
  return -2.0 * PI * (double)(ivk1 * n1 + ivk2 * n1 + ivk2 * n2) / (double)p->numk;
}

/* Solve, HartreeFock.f90:402-483: OpenMP over k; zFockBulk is returned
 * unnormalised (the reference divides by Nk in Main.f90:510-511). */
int orc_solve(const orc_params* p, zc* zFockBulk, const zc* zFock, const double* Potential,
              double* Bands, zc* zEig, const double* C, const double* MomVal,
              const i64* nMomComp, const i64* u1, const i64* u2, const double* tn,
              int nthreads) {
  i64 n = p->ndim, nnc = p->nnc, numk = p->numk, Nk = numk * numk;
  i64 nL = p->nLower, nU = p->nUpper, nb = nU - nL + 1;
  size_t nn = (size_t)n * n;
  memset(zFockBulk, 0, sizeof(zc) * nn * nnc);
  int status = 0;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
  {
    zc* H = (zc*)malloc(sizeof(zc) * nn);
    zc* T = (zc*)malloc(sizeof(zc) * nn);
    zc* zP = (zc*)malloc(sizeof(zc) * n);
    double* E = (double*)malloc(sizeof(double) * n);
    zc* acc = (zc*)calloc(nn, sizeof(zc));
    zc* accc = (zc*)calloc(nn * (nnc - 1 > 0 ? nnc - 1 : 1), sizeof(zc));
#pragma omp for schedule(dynamic, 1)
    for (i64 ic = 0; ic < Nk; ++ic) {
      i64 i1 = nMomComp[ic], i2 = nMomComp[ic + Nk];
      double vk[2] = {MomVal[i1 + numk * i2], MomVal[i1 + numk * i2 + Nk]};
      orc_hamiltonian_hf(p, H, C, Potential, zFock, u1, u2, vk, tn);
      int info = orc_zdiagred(H, n, E, 1, nU);
      if (info) {
#pragma omp atomic write
        status = info;
      }
      for (i64 b = nL; b <= nU; ++b) {
        Bands[(b - nL) + nb * ic] = E[b - 1];
        memcpy(zEig + n * ((b - nL) + nb * ic), H + n * (b - 1), sizeof(zc) * n);
      }
      memset(T, 0, sizeof(zc) * nn);
      for (i64 b = 1; b <= nL - 1; ++b) {
        for (i64 i = 0; i < n; ++i) {
          double ph = (vk[0] * C[i] + vk[1] * C[i + n]) * (double)p->fphase;
          zP[i] = conj(H[i + n * (b - 1)] * cexp(I * ph));
        }
        vector_mul(zP, T, (int)n, 1.0);
      }
      for (size_t q = 0; q < nn; ++q) acc[q] += T[q];
      for (i64 c = 1; c < nnc; ++c) {
        zc f = cexp(I * cell_phase_arg(p, i1, i2, u1[c], u2[c]));
        zc* dst = accc + nn * (c - 1);
        for (size_t q = 0; q < nn; ++q) dst[q] += T[q] * f;
      }
    }
#pragma omp critical
    {
      for (size_t q = 0; q < nn; ++q) zFockBulk[q] += acc[q];
      for (i64 c = 1; c < nnc; ++c)
        for (size_t q = 0; q < nn; ++q) zFockBulk[q + nn * c] += accc[q + nn * (c - 1)];
    }
    free(H); free(T); free(zP); free(E); free(acc); free(accc);
  }
  return status;
}

/* ------------------------------------------------------------------ sorting / Fermi */

/* dsort_lists, TightBinding.f90:984-1027 (heap sort on column 1, rows carry a
 * second column; here the second column is the flat index). */
static void heap_sort2(double* key, double* val, i64 N) {
  if (N < 2) return;
  i64 l = N / 2 + 1, ir = N, i, j;
  double ck, cv;
  for (;;) {
    if (l > 1) {
      --l; ck = key[l - 1]; cv = val[l - 1];
    } else {
      ck = key[ir - 1]; cv = val[ir - 1];
      key[ir - 1] = key[0]; val[ir - 1] = val[0];
      --ir;
      if (ir == 1) { key[0] = ck; val[0] = cv; return; }
    }
    i = l; j = 2 * l;
    while (j <= ir) {
      if (j < ir && key[j - 1] < key[j]) ++j;
      if (ck < key[j - 1]) {
        key[i - 1] = key[j - 1]; val[i - 1] = val[j - 1];
        i = j; j = 2 * j;
      } else {
        j = ir + 1;
      }
    }
    key[i - 1] = ck; val[i - 1] = cv;
  }
}

/* SortEnergies_1Spin, HartreeFock.f90:885-925.  sorted_index (optional)
 * receives the 1-based flat index icount = (b-1)*Nk + k of each sorted state. */
void orc_sort_energies(const orc_params* p, const double* Bands, const zc* zEig,
                       const i64* nMomComp, double* Sorted, zc* zSorted, i64* nSortedMom,
                       i64* sorted_index) {
  i64 n = p->ndim, Nk = p->numk * p->numk, nb = p->nUpper - p->nLower + 1, N = nb * Nk;
  double* key = (double*)malloc(sizeof(double) * N);
  double* val = (double*)malloc(sizeof(double) * N);
  for (i64 ic = 1; ic <= N; ++ic) {
    i64 indev = (ic - 1) / Nk + 1;
    key[ic - 1] = Bands[(indev - 1) + nb * (ic - (indev - 1) * Nk - 1)];
    val[ic - 1] = (double)ic;
  }
  heap_sort2(key, val, N);
  for (i64 ic = 1; ic <= N; ++ic) {
    i64 src = (i64)llround(val[ic - 1]);
    i64 indev = (src - 1) / Nk + 1;
    i64 k = src - (indev - 1) * Nk;
    Sorted[ic - 1] = key[ic - 1];
    if (zSorted) memcpy(zSorted + n * (ic - 1), zEig + n * ((indev - 1) + nb * (k - 1)), sizeof(zc) * n);
    if (nSortedMom) {
      nSortedMom[ic - 1] = nMomComp[k - 1];
      nSortedMom[ic - 1 + N] = nMomComp[k - 1 + Nk];
    }
    if (sorted_index) sorted_index[ic - 1] = src;
  }
  free(key); free(val);
}

/* Fortran nint: round half away from zero. */
static i64 f_nint(double x) { return (i64)(x >= 0 ? floor(x + 0.5) : -floor(-x + 0.5)); }

/* Fermi block, Main.f90:400-412 (numS = 1). */
void orc_fermi(const orc_params* p, const double* S, i64 N, i64* nF, double* EFermi,
               i64* nOcc, i64* nPart, double* Deg) {
  i64 Nk = p->numk * p->numk;
  i64 NeutralityPoint = (p->ndim / 2 - p->nLower) + 1;
  *nF = f_nint((double)(Nk * NeutralityPoint) + (double)Nk / 2.0 * p->nfilling);
  *EFermi = (S[*nF - 1] + S[*nF]) / 2.0;
  i64 o = 0, q = 0;
  for (i64 ic = 0; ic < N; ++ic) {
    if ((S[ic] - *EFermi) < -p->energy_tol) ++o;
    else if (fabs(S[ic] - *EFermi) < p->energy_tol) { ++o; ++q; }
  }
  *nOcc = o; *nPart = q;
  *Deg = 1.0 / (double)q * (double)(*nF + q - o);
}

/* Fermi-Dirac occupations of the sorted window (extension): weights f_s with
 * sum f_s = nF, solved for the chemical potential by bisection. */
double orc_fermi_dirac(const double* S, i64 N, i64 nF, double kT, double* w) {
  double lo = S[0] - 50 * kT - 1.0, hi = S[N - 1] + 50 * kT + 1.0, mu = 0;
  for (int it = 0; it < 200; ++it) {
    mu = 0.5 * (lo + hi);
    double s = 0;
    for (i64 i = 0; i < N; ++i) s += 1.0 / (1.0 + exp((S[i] - mu) / kT));
    if (s > (double)nF) hi = mu; else lo = mu;
  }
  for (i64 i = 0; i < N; ++i) w[i] = 1.0 / (1.0 + exp((S[i] - mu) / kT));
  return mu;
}

/* GetFock, HartreeFock.f90:487-560.  occ_mode 0 reproduces the reference's
 * degenerate-shell weight zdeg = cmplx(DegFactor, dp) (HartreeFock.f90:529),
 * whose second argument is the imaginary part: zdeg = (float(DegFactor), 8).
 * occ_mode 1 uses the real weight DegFactor; occ_mode 2 uses weights[s] for
 * every sorted state s < count. zFock is accumulated into (not cleared). */
void orc_getfock(const orc_params* p, zc* zFock, const zc* zSorted, const i64* nSortedMom,
                 double Deg, i64 nOcc, i64 nPart, const double* C, const double* MomVal,
                 const i64* u1, const i64* u2, int occ_mode, const double* weights, i64 count) {
  i64 n = p->ndim, nnc = p->nnc, numk = p->numk, Nk = numk * numk;
  i64 N = (p->nUpper - p->nLower + 1) * Nk;
  size_t nn = (size_t)n * n;
  zc zdeg = occ_mode == 0 ? ((double)(float)Deg + 8.0 * I) : (zc)Deg;
  i64 first = 0, last;
  if (occ_mode == 2) last = count; else last = nOcc;
  zc* zP = (zc*)malloc(sizeof(zc) * n);
  for (i64 s = first; s < last && s < N; ++s) {
    i64 i1 = nSortedMom[s], i2 = nSortedMom[s + N];
    double vk[2] = {MomVal[i1 + numk * i2], MomVal[i1 + numk * i2 + Nk]};
    for (i64 i = 0; i < n; ++i) {
      double ph = (vk[0] * C[i] + vk[1] * C[i + n]) * (double)p->fphase;
      zP[i] = conj(zSorted[i + n * s] * cexp(I * ph));
    }
    zc wgt;
    if (occ_mode == 2) wgt = weights[s];
    else wgt = (s < nOcc - nPart) ? (zc)1.0 : zdeg;
    if (occ_mode == 2 && weights[s] == 0.0) continue;
    for (i64 c = 0; c < nnc; ++c) {
      zc f = cexp(I * cell_phase_arg(p, i1, i2, u1[c], u2[c])) * wgt;
      vector_mul(zP, zFock + nn * c, (int)n, f);
    }
  }
  free(zP);
}

/* ------------------------------------------------------------------ C3-reduced k-set */

/* C3_RelatedPoints, Geometry.f90:1172-1223: perm[i] (0-based) is the site j of
 * the same layer and sublattice half that sits at C3 r_i (rotation by +120
 * degrees about the origin) within 1e-3; the last site of each half (the two
 * cell corners of each layer) maps to itself.  Returns the number of sites
 * assigned (ndim for a C3-symmetric cell; the reference prints an error
 * otherwise and leaves unmatched entries unset, here they are -1). */
i64 orc_c3_pairs(const double* C, i64 ndim, i64 nlayers, i64* perm) {
  const double h3 = 0.5 * sqrt(3.0);
  i64 per = ndim / nlayers, half = per / 2, cnt = 0;
  for (i64 i = 0; i < ndim; ++i) perm[i] = -1;
  for (i64 l = 0; l < nlayers; ++l)
    for (int hf = 0; hf < 2; ++hf) {
      i64 lo = l * per + hf * half, hi = lo + half;
      for (i64 i = lo; i < hi; ++i) {
        double x = C[i], y = C[i + ndim];
        double x3 = -0.5 * x - h3 * y, y3 = h3 * x - 0.5 * y;
        for (i64 j = lo; j < hi; ++j) {
          double dx = C[j] - x3, dy = C[j + ndim] - y3;
          if (sqrt(dx * dx + dy * dy) < 1e-3) { perm[i] = j; ++cnt; }
        }
      }
    }
  for (i64 l = 0; l < nlayers; ++l) {
    perm[l * per + half - 1] = l * per + half - 1;
    perm[l * per + per - 1] = l * per + per - 1;
    cnt += 2;
  }
  return cnt;
}

/* Corner-site factors of the C3 images.  For layer rotation sign rot, the
 * first corner (last site of the first half) carries e^{i rot a} and the second
 * (last site of the layer) e^{-i rot a}; every other site carries 1.  The
 * reference writes these as row/column scalings of the rotated matrices
 * (HartreeFock.f90:300-330, 344-372) and as zphase of the rotated eigenvectors
 * (HartreeFock.f90:225-275). */
static void c3_corner(const orc_params* p, const i64* rot, double a, zc* f) {
  i64 n = p->ndim, per = n / p->nlayers, half = per / 2;
  for (i64 i = 0; i < n; ++i) f[i] = 1.0;
  for (i64 l = 0; l < p->nlayers; ++l) {
    double r = (double)rot[l];
    if (rot[l] != 1 && rot[l] != -1) continue;
    f[l * per + half - 1] = cexp(I * r * a);
    f[l * per + per - 1] = cexp(-I * r * a);
  }
}

/* Flat momentum index ivk1*numk + ivk2 (0-based) of the two C3 partners of k,
 * HartreeFock.f90:152-165: rel[0] = (-i1-i2, i1), rel[1] = (i2, -i1-i2) mod numk. */
static void c3_related(i64 numk, i64 i1, i64 i2, i64* rel) {
  i64 m = ((-i1 - i2) % numk + numk) % numk;
  rel[0] = m * numk + (i1 % numk);
  rel[1] = (i2 % numk) * numk + m;
}

/* The C3-invariant momenta Gamma, K, K' (HartreeFock.f90:218, 377, 594):
 * flat indices 0, numk/3 (numk+1), 2 numk/3 (numk+1); numk % 3 == 0. */
static int c3_invariant(i64 numk, i64 k) {
  return k == 0 || k == numk / 3 * (numk + 1) || k == 2 * numk / 3 * (numk + 1);
}

/* T1(i,j) = T(perm i, perm j) f_i conj(f_j): the C3 image of a density block
 * (the reference's row/column scalings, with f the row factor). */
static void c3_image(i64 n, const i64* perm, const zc* f, const zc* T, zc* out) {
  for (i64 j = 0; j < n; ++j)
    for (i64 i = 0; i < n; ++i)
      out[i + n * j] = T[perm[i] + n * perm[j]] * f[i] * conj(f[j]);
}

/* Solve_C3, HartreeFock.f90:131-398.  Diagonalises H(k) only at one momentum
 * of each C3 orbit (the first in flat order), copies its window bands to the
 * two partners, writes the partners' window eigenvectors as permuted and
 * corner-phased copies, and adds the bulk density of the orbit as
 * T phi_c(k) + C3[T] phi_c(k'') + C3^2[T] phi_c(k'), divided by 3 at the
 * invariant momenta.  Returns -2 when numk % 3 != 0 (the reference's
 * invariant-point test is only right then), else zheevr's info. */
int orc_solve_c3(const orc_params* p, const i64* rot, const i64* perm, zc* zFockBulk,
                 const zc* zFock, const double* Potential, double* Bands, zc* zEig,
                 const double* C, const double* MomVal, const i64* nMomComp, const i64* u1,
                 const i64* u2, const double* tn, int nthreads) {
  i64 n = p->ndim, nnc = p->nnc, numk = p->numk, Nk = numk * numk;
  i64 nL = p->nLower, nU = p->nUpper, nb = nU - nL + 1;
  size_t nn = (size_t)n * n;
  if (numk % 3 != 0) return -2;
  for (i64 k = 0; k < Nk; ++k)
    if (nMomComp[k] * numk + nMomComp[k + Nk] != k) return -2;
  /* representatives, HartreeFock.f90:168-189 */
  i64* reps = (i64*)malloc(sizeof(i64) * Nk);
  char* incl = (char*)calloc(Nk, 1);
  i64 nrep = 0;
  for (i64 k = 0; k < Nk; ++k) {
    if (incl[k]) continue;
    i64 rel[2];
    c3_related(numk, nMomComp[k], nMomComp[k + Nk], rel);
    reps[nrep++] = k;
    incl[k] = incl[rel[0]] = incl[rel[1]] = 1;
  }
  memset(zFockBulk, 0, sizeof(zc) * nn * nnc);
  int status = 0;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
  {
    zc* H = (zc*)malloc(sizeof(zc) * nn);
    zc* T = (zc*)malloc(sizeof(zc) * nn);
    zc* R1 = (zc*)malloc(sizeof(zc) * nn);
    zc* R2 = (zc*)malloc(sizeof(zc) * nn);
    zc* zP = (zc*)malloc(sizeof(zc) * n);
    zc* f1 = (zc*)malloc(sizeof(zc) * n);
    zc* f2 = (zc*)malloc(sizeof(zc) * n);
    double* E = (double*)malloc(sizeof(double) * n);
    zc* acc = (zc*)calloc(nn * nnc, sizeof(zc));
#pragma omp for schedule(dynamic, 1)
    for (i64 r = 0; r < nrep; ++r) {
      i64 k = reps[r], i1 = nMomComp[k], i2 = nMomComp[k + Nk];
      i64 rel[2];
      c3_related(numk, i1, i2, rel);
      const int inv = c3_invariant(numk, k);
      double vk[2] = {MomVal[i1 + numk * i2], MomVal[i1 + numk * i2 + Nk]};
      orc_hamiltonian_hf(p, H, C, Potential, zFock, u1, u2, vk, tn);
      int info = orc_zdiagred(H, n, E, 1, nU);
      if (info) {
#pragma omp atomic write
        status = info;
      }
      for (i64 b = nL; b <= nU; ++b)
        for (int q = 0; q < 3; ++q) Bands[(b - nL) + nb * (q == 0 ? k : rel[q - 1])] = E[b - 1];
      /* window eigenvectors: C3^2 image at rel[0], C3 image at rel[1]
       * (HartreeFock.f90:212-277, fphase = 0) */
      i64 k2 = rel[0], k3 = rel[1];
      c3_corner(p, rot, 2.0 * PI * (double)(nMomComp[k2] + nMomComp[k2 + Nk]) / (double)numk, f2);
      c3_corner(p, rot, 2.0 * PI * (double)nMomComp[k3 + Nk] / (double)numk, f1);
      for (i64 b = nL; b <= nU; ++b) {
        const zc* v = H + n * (b - 1);
        memcpy(zEig + n * ((b - nL) + nb * k), v, sizeof(zc) * n);
        if (inv) continue;
        zc* w2 = zEig + n * ((b - nL) + nb * k2);
        zc* w1 = zEig + n * ((b - nL) + nb * k3);
        for (i64 i = 0; i < n; ++i) {
          w2[i] = v[perm[perm[i]]] * f2[i];
          w1[i] = v[perm[i]] * f1[i];
        }
      }
      /* bulk block of the representative, HartreeFock.f90:280-295 */
      memset(T, 0, sizeof(zc) * nn);
      for (i64 b = 1; b <= nL - 1; ++b) {
        for (i64 i = 0; i < n; ++i) {
          double ph = (vk[0] * C[i] + vk[1] * C[i + n]) * (double)p->fphase;
          zP[i] = conj(H[i + n * (b - 1)] * cexp(I * ph));
        }
        vector_mul(zP, T, (int)n, 1.0);
      }
      /* C3 image (first-corner rows e^{i rot t1}, t1 = 2 pi (i1+i2)/numk) and
       * its C3 image (first-corner rows e^{-i rot t2}, t2 = 2 pi i1/numk),
       * HartreeFock.f90:297-375 */
      c3_corner(p, rot, 2.0 * PI * (double)(i1 + i2) / (double)numk, f1);
      c3_image(n, perm, f1, T, R1);
      c3_corner(p, rot, -2.0 * PI * (double)i1 / (double)numk, f2);
      c3_image(n, perm, f2, R1, R2);
      const double w = inv ? 1.0 / 3.0 : 1.0;
      for (i64 c = 0; c < nnc; ++c) {
        zc ph0 = c == 0 ? 1.0 : cexp(I * cell_phase_arg(p, i1, i2, u1[c], u2[c]));
        zc ph1 = c == 0 ? 1.0 : cexp(I * cell_phase_arg(p, nMomComp[k3], nMomComp[k3 + Nk], u1[c], u2[c]));
        zc ph2 = c == 0 ? 1.0 : cexp(I * cell_phase_arg(p, nMomComp[k2], nMomComp[k2 + Nk], u1[c], u2[c]));
        zc* dst = acc + nn * c;
        for (size_t q = 0; q < nn; ++q) dst[q] += w * (T[q] * ph0 + R1[q] * ph1 + R2[q] * ph2);
      }
    }
#pragma omp critical
    for (size_t q = 0; q < nn * nnc; ++q) zFockBulk[q] += acc[q];
    free(H); free(T); free(R1); free(R2); free(zP); free(f1); free(f2); free(E); free(acc);
  }
  free(reps); free(incl);
  return status;
}

/* GetFock_C3, HartreeFock.f90:564-881.  States at C3-non-invariant momenta
 * add as in GetFock (their eigenvectors are already the C3 images written by
 * Solve_C3); a state at an invariant momentum adds (T + C3[T] + C3^2[T]) / 3
 * to every cell with its own cell phase.  Here the images are formed on the
 * state vector (zP1 = f1 zP(perm), zP2 = f2 zP1(perm)), which is the same
 * product as the reference's row/column scaling of the permuted block.
 * Weights as in orc_getfock. */
void orc_getfock_c3(const orc_params* p, const i64* rot, const i64* perm, zc* zFock,
                    const zc* zSorted, const i64* nSortedMom, double Deg, i64 nOcc, i64 nPart,
                    const double* C, const double* MomVal, const i64* u1, const i64* u2,
                    int occ_mode, const double* weights, i64 count) {
  i64 n = p->ndim, nnc = p->nnc, numk = p->numk, Nk = numk * numk;
  i64 N = (p->nUpper - p->nLower + 1) * Nk;
  size_t nn = (size_t)n * n;
  zc zdeg = occ_mode == 0 ? ((double)(float)Deg + 8.0 * I) : (zc)Deg;
  i64 last = occ_mode == 2 ? count : nOcc;
  zc* zP = (zc*)malloc(sizeof(zc) * 3 * n);
  zc* f = (zc*)malloc(sizeof(zc) * n);
  for (i64 s = 0; s < last && s < N; ++s) {
    i64 i1 = nSortedMom[s], i2 = nSortedMom[s + N];
    double vk[2] = {MomVal[i1 + numk * i2], MomVal[i1 + numk * i2 + Nk]};
    for (i64 i = 0; i < n; ++i) {
      double ph = (vk[0] * C[i] + vk[1] * C[i + n]) * (double)p->fphase;
      zP[i] = conj(zSorted[i + n * s] * cexp(I * ph));
    }
    zc wgt;
    if (occ_mode == 2) wgt = weights[s];
    else wgt = (s < nOcc - nPart) ? (zc)1.0 : zdeg;
    if (occ_mode == 2 && weights[s] == 0.0) continue;
    int nvec = 1;
    if (c3_invariant(numk, i1 * numk + i2)) {
      nvec = 3;
      wgt /= 3.0;
      c3_corner(p, rot, 2.0 * PI * (double)(i1 + i2) / (double)numk, f);
      for (i64 i = 0; i < n; ++i) zP[n + i] = zP[perm[i]] * f[i];
      c3_corner(p, rot, -2.0 * PI * (double)i1 / (double)numk, f);
      for (i64 i = 0; i < n; ++i) zP[2 * n + i] = zP[n + perm[i]] * f[i];
    }
    for (i64 c = 0; c < nnc; ++c) {
      zc ph = cexp(I * cell_phase_arg(p, i1, i2, u1[c], u2[c])) * wgt;
      for (int q = 0; q < nvec; ++q) vector_mul(zP + n * q, zFock + nn * c, (int)n, ph);
    }
  }
  free(zP); free(f);
}

/* ------------------------------------------------------------------ T / C2 / C2T */

/* C2_RelatedPoints, Geometry.f90:1140-1167: within each layer, site i of the
 * first sublattice half and site j of the second are partners when
 * r_i = -r_j in the plane (within 1e-5); perm[i] = j and perm[j] = i
 * (0-based, -1 when unmatched).  Returns the number of pairs (ndim/2 for a
 * C2-symmetric cell). */
i64 orc_c2_pairs(const double* C, i64 ndim, i64 nlayers, i64* perm) {
  i64 per = ndim / nlayers, half = per / 2, cnt = 0;
  for (i64 i = 0; i < ndim; ++i) perm[i] = -1;
  for (i64 l = 0; l < nlayers; ++l)
    for (i64 i = l * per; i < l * per + half; ++i)
      for (i64 j = l * per + half; j < (l + 1) * per; ++j) {
        double x = C[i] + C[j], y = C[i + ndim] + C[j + ndim];
        if (sqrt(x * x + y * y) < 1e-5) { perm[i] = j; perm[j] = i; ++cnt; }
      }
  return cnt;
}

/* The symmetrisations of Main.f90:514-534, in that order, on every stored
 * cell of rho: sym & 1 (nenforceT) keeps the real part; sym & 2 (nenforceC2)
 * rho <- (rho + conj(rho(p, p))^T) / 2; sym & 4 (nenforceC2T)
 * rho <- (rho + rho(p, p)^T) / 2, p = C2 partners. */
void orc_symmetrize(const orc_params* p, zc* rho, int sym, const i64* c2) {
  i64 n = p->ndim;
  size_t nn = (size_t)n * n;
  if (sym & 1)
    for (size_t q = 0; q < nn * p->nnc; ++q) rho[q] = creal(rho[q]);
  zc* t = (zc*)malloc(sizeof(zc) * nn);
  for (int f = 2; f <= 4; f *= 2) {
    if (!(sym & f)) continue;
    for (i64 m = 0; m < p->nnc; ++m) {
      zc* r = rho + nn * m;
      memcpy(t, r, sizeof(zc) * nn);
      for (i64 j = 0; j < n; ++j)
        for (i64 i = 0; i < n; ++i) {
          zc o = t[c2[j] + n * c2[i]];
          r[i + n * j] = 0.5 * (t[i + n * j] + (f == 2 ? conj(o) : o));
        }
    }
  }
  free(t);
}

/* ------------------------------------------------------------------ valley projection */

/* ftvalley, TightBinding.f90:235-270: i sign(cos 3 phi) / sqrt(27) for sites
 * 0.2 < |r_i - r_j| < 1.2 apart in the plane, phi the angle of
 * R (r_i - r_j) (R the 2x2 rotation given row-major), else 0. */
static zc ftvalley(const double* ri, const double* rj, const double* R) {
  double dx = ri[0] - rj[0], dy = ri[1] - rj[1];
  double rij = sqrt(dx * dx + dy * dy);
  if (!(rij < 1.2 && rij > 0.2)) return 0.0;
  double x = R[0] * dx + R[1] * dy, y = R[2] * dx + R[3] * dy, phi;
  if (x > 0) phi = atan(y / x);
  else if (x < 0 && y > 0) phi = atan(y / x) + PI;
  else phi = -PI + atan(y / x);
  if (x == 0) phi = y > 0 ? PI / 2.0 : -PI / 2.0;
  return I * copysign(1.0, cos(3.0 * phi)) / sqrt(27.0);
}

/* ValleyPhase, TightBinding.f90:60-165, for fphase = 0: V(i,j) over the
 * sites i, j of one sublattice half of one layer, home cell plus both
 * images in every other stored cell, with R = RotMatrix for layers rotated
 * +1 and its transpose for -1 (other layers contribute nothing).
 * RotMatrix = [[c, s], [-s, c]], c, s = cos, sin(acos(cs)/2) (Main.f90:133). */
static void valley_phase(const orc_params* p, const i64* rot, const double* C, const double* tn,
                         const i64* u1, const i64* u2, const double* vk, zc* V) {
  i64 n = p->ndim, per = n / p->nlayers, half = per / 2;
  double aM = (double)(n / (2 * p->nlayers));  /* 3 ntheta^2 + 3 ntheta + 1 */
  double cs = 1.0 - 1.0 / (2.0 * aM);
  double c = cos(0.5 * acos(cs)), sn = sin(0.5 * acos(cs));
  const double Rp[4] = {c, sn, -sn, c}, Rm[4] = {c, -sn, sn, c};
  memset(V, 0, sizeof(zc) * n * n);
  for (i64 l = 0; l < p->nlayers; ++l) {
    const double* R = rot[l] == -1 ? Rm : (rot[l] == 1 ? Rp : NULL);
    if (!R) continue;
    for (int hf = 0; hf < 2; ++hf) {
      i64 lo = l * per + hf * half, hi = lo + half;
      for (i64 i = lo; i < hi; ++i)
        for (i64 j = lo; j < hi; ++j) {
          double ri[2] = {C[i], C[i + n]}, rj[2] = {C[j], C[j + n]};
          zc v = ftvalley(ri, rj, R);
          for (i64 cc = 1; cc < p->nnc; ++cc) {
            double Tx = (double)u1[cc] * tn[0] + (double)u2[cc] * tn[2];
            double Ty = (double)u1[cc] * tn[1] + (double)u2[cc] * tn[3];
            double ph = vk[0] * Tx + vk[1] * Ty;
            double rjs[2] = {rj[0] - Tx, rj[1] - Ty}, ris[2] = {ri[0] - Tx, ri[1] - Ty};
            v += ftvalley(ri, rjs, R) * cexp(-I * ph);
            v += ftvalley(ris, rj, R) * cexp(I * ph);
          }
          V[i + n * j] += v;
        }
    }
  }
}

/* valley_phase for one momentum (exported for the tests). */
void orc_valley_phase(const orc_params* p, const i64* rot, const double* C, const double* tn,
                      const i64* u1, const i64* u2, const double* vk, zc* V) {
  valley_phase(p, rot, C, tn, u1, u2, vk, V);
}

/* ValleyTransform, TightBinding.f90:170-211: per k, P = C^H V C over the
 * window columns, zheev('V', 'L') of P (zdiagFull, LapackRoutines.f90:498-533),
 * S = U sign(Lambda) U^H (unitarize, LapackRoutines.f90:865-893; sign(1, x)
 * keeps the sign of a negative zero), C <- C S, each product as the
 * reference's BLAS call forms it (the SCF amplifies their rounding
 * differences).  zEig: n x nb x Nk. */
void orc_valley_transform(const orc_params* p, const i64* rot, const double* C, const double* tn,
                          const double* MomVal, const i64* nMomComp, const i64* u1, const i64* u2,
                          zc* zEig) {
  i64 n = p->ndim, numk = p->numk, Nk = numk * numk, nb = p->nUpper - p->nLower + 1;
  int ni = (int)n, mb = (int)nb;
  zc* V = (zc*)malloc(sizeof(zc) * n * n);
  zc* D = (zc*)malloc(sizeof(zc) * n * nb);
  zc* Cc = (zc*)malloc(sizeof(zc) * n * nb);
  zc* P = (zc*)malloc(sizeof(zc) * nb * nb);
  zc* A = (zc*)malloc(sizeof(zc) * nb * nb);
  zc* Bd = (zc*)malloc(sizeof(zc) * nb * nb);
  zc* S = (zc*)malloc(sizeof(zc) * nb * nb);
  zc* T = (zc*)malloc(sizeof(zc) * n * nb);
  double* w = (double*)malloc(sizeof(double) * nb);
  double* rwork = (double*)malloc(sizeof(double) * 3 * nb);
  const zc one = 1.0, zero = 0.0;
  for (i64 k = 0; k < Nk; ++k) {
    i64 i1 = nMomComp[k], i2 = nMomComp[k + Nk];
    double vk[2] = {MomVal[i1 + numk * i2], MomVal[i1 + numk * i2 + Nk]};
    zc* Ck = zEig + n * nb * k;
    valley_phase(p, rot, C, tn, u1, u2, vk, V);
    /* matrix_elements (LapackRoutines.f90:257-265): D = V C, P = conjg(C)^T D */
    zgemm_("N", "N", &ni, &mb, &ni, &one, V, &ni, Ck, &ni, &zero, D, &ni, 1, 1);
    for (i64 q = 0; q < n * nb; ++q) Cc[q] = conj(Ck[q]);
    zgemm_("T", "N", &mb, &mb, &ni, &one, Cc, &ni, D, &ni, &zero, P, &mb, 1, 1);
    /* unitarize: zheev('V', 'L'), B = diag(sign w), A = U^H, S = conjg(A)^T (B A) */
    int lwork = -1, info;
    zc wq;
    zheev_("V", "L", &mb, P, &mb, w, &wq, &lwork, rwork, &info, 1, 1);
    lwork = (int)creal(wq);
    zc* work = (zc*)malloc(sizeof(zc) * (lwork > 1 ? lwork : 1));
    zheev_("V", "L", &mb, P, &mb, w, work, &lwork, rwork, &info, 1, 1);
    free(work);
    for (i64 b = 0; b < nb; ++b)
      for (i64 a = 0; a < nb; ++a) {
        A[a + nb * b] = conj(P[b + nb * a]);
        Bd[a + nb * b] = a == b ? copysign(1.0, w[a]) : 0.0;
      }
    zgemm_("N", "N", &mb, &mb, &mb, &one, Bd, &mb, A, &mb, &zero, Cc, &mb, 1, 1);
    for (i64 q = 0; q < nb * nb; ++q) A[q] = conj(A[q]);
    zgemm_("T", "N", &mb, &mb, &mb, &one, A, &mb, Cc, &mb, &zero, S, &mb, 1, 1);
    /* matrix_mul: C S */
    zgemm_("N", "N", &ni, &mb, &mb, &one, Ck, &ni, S, &mb, &zero, T, &ni, 1, 1);
    memcpy(Ck, T, sizeof(zc) * n * nb);
  }
  free(V); free(D); free(Cc); free(P); free(A); free(Bd); free(S); free(T); free(w); free(rwork);
}

/* ------------------------------------------------------------------ energies */

/* GetKineticEnergy, HartreeFock.f90:1012-1059. */
double orc_kinetic_energy(const orc_params* p, const zc* zF, const double* C, const i64* u1,
                          const i64* u2, const double* tn) {
  i64 n = p->ndim, nnc = p->nnc;
  size_t nn = (size_t)n * n;
  const double* t1 = tn; const double* t2 = tn + 2;
  double sr = 0, si = 0;
#pragma omp parallel for reduction(+ : sr, si) schedule(dynamic, 8)
  for (i64 i = 0; i < n; ++i) {
    double ri[3], rj[3], rs[3];
    site(C, n, i, ri);
    zc z = 0;
    for (i64 j = 0; j < n; ++j) {
      site(C, n, j, rj);
      z += orc_ft(p, ri, rj) * zF[i + n * j];
      for (i64 c = 1; c < nnc; ++c) {
        double Tx = u1[c] * t1[0] + u2[c] * t2[0], Ty = u1[c] * t1[1] + u2[c] * t2[1];
        rs[0] = rj[0] - Tx; rs[1] = rj[1] - Ty; rs[2] = rj[2];
        z += orc_ft(p, ri, rs) * zF[i + n * j + nn * c];
        rs[0] = rj[0] + Tx; rs[1] = rj[1] + Ty;
        z += orc_ft(p, ri, rs) * conj(zF[j + n * i + nn * c]);
      }
    }
    sr += creal(z); si += cimag(z);
  }
  for (i64 i = 0; i < n; ++i) sr += creal(zF[i + n * i] * layer_bias(p, i));
  return sr;
}

/* GetFockEnergy, HartreeFock.f90:1098-1145. */
double orc_fock_energy(const orc_params* p, const zc* zF, const double* C, const i64* u1,
                       const i64* u2, const double* tn) {
  i64 n = p->ndim, nnc = p->nnc;
  size_t nn = (size_t)n * n;
  const double* t1 = tn; const double* t2 = tn + 2;
  double sr = 0;
#pragma omp parallel for reduction(+ : sr) schedule(dynamic, 8)
  for (i64 i = 0; i < n; ++i) {
    double ri[3], rj[3], rs[3];
    site(C, n, i, ri);
    double z = 0;
    for (i64 j = 0; j < n; ++j) {
      site(C, n, j, rj);
      zc a = zF[i + n * j];
      z += -0.5 * orc_fv(p, ri, rj) * creal(conj(a) * a);
      for (i64 c = 1; c < nnc; ++c) {
        double Tx = u1[c] * t1[0] + u2[c] * t2[0], Ty = u1[c] * t1[1] + u2[c] * t2[1];
        zc b = zF[i + n * j + nn * c], d = zF[j + n * i + nn * c];
        rs[0] = rj[0] - Tx; rs[1] = rj[1] - Ty; rs[2] = rj[2];
        z += -0.5 * orc_fv(p, ri, rs) * creal(conj(b) * b);
        rs[0] = rj[0] + Tx; rs[1] = rj[1] + Ty;
        z += -0.5 * orc_fv(p, ri, rs) * creal(d * conj(d));
      }
    }
    sr += z;
  }
  return sr;
}

/* GetHartreeHubbardEnergy, HartreeFock.f90:1066-1092. */
void orc_hartree_hubbard(const orc_params* p, const double* L, const double* d1,
                         const double* d2, double* EH, double* EU) {
  i64 n = p->ndim;
  double h = 0;
#pragma omp parallel for reduction(+ : h)
  for (i64 a = 0; a < n; ++a)
    for (i64 b = 0; b < n; ++b) h += 0.5 * (d1[b] + d2[b]) * (d1[a] + d2[a]) * L[a + n * b];
  double u = 0;
  for (i64 a = 0; a < n; ++a) u += d1[a] * d2[a];
  *EH = h; *EU = u;
}

/* All four energies of rho with densities dens (= Re diag rho - 1/2). */
void orc_energies(const orc_params* p, const zc* zF, const double* dens, const double* C,
                  const double* L, const i64* u1, const i64* u2, const double* tn, double* out) {
  out[0] = orc_kinetic_energy(p, zF, C, u1, u2, tn);
  out[1] = orc_fock_energy(p, zF, C, u1, u2, tn);
  orc_hartree_hubbard(p, L, dens, dens, &out[2], &out[3]);
}

/* TotalEnergy, Main.f90:312 (numS = 1). */
double orc_total_energy(const orc_params* p, const double* e) {
  return p->alpha * e[1] * 2.0 + e[0] * 2.0 + p->U * e[3] + p->alphaH * e[2];
}

/* Potential, Main.f90:577-585 (numS = 1). */
void orc_potential(const orc_params* p, const double* L, const double* Density, double* V) {
  i64 n = p->ndim;
#pragma omp parallel for
  for (i64 a = 0; a < n; ++a) {
    double s = 0;
    for (i64 b = 0; b < n; ++b) s += 2.0 * (Density[b] - 0.5) * p->alphaH * L[a + n * b];
    V[a] = s;
  }
  for (i64 a = 0; a < n; ++a) V[a] += p->U * (Density[a] - 0.5);
}

/* OptimalStep, HartreeFock.f90:1148-1227 (numS = 1).  e_inout holds the
 * energies (Ekin, EF, EH, EU) of rho_in on entry and those of the mixed
 * density on exit; returns Step. */
double orc_optimal_step(const orc_params* p, double* e_inout, const zc* rhoNew, const zc* rhoIn,
                        const double* Dens, const double* DensIn, const double* C,
                        const double* L, const i64* u1, const i64* u2, const double* tn) {
  i64 n = p->ndim;
  size_t tot = (size_t)n * n * p->nnc;
  double c0 = orc_total_energy(p, e_inout);
  double* dt = (double*)malloc(sizeof(double) * n);
  zc* tmp = (zc*)malloc(sizeof(zc) * tot);
  double e[4];
  for (i64 a = 0; a < n; ++a) dt[a] = Dens[a] - 0.5;
  orc_energies(p, rhoNew, dt, C, L, u1, u2, tn, e);
  double E1 = orc_total_energy(p, e);
  for (i64 a = 0; a < n; ++a) dt[a] = 0.5 * (Dens[a] + DensIn[a]) - 0.5;
  for (size_t q = 0; q < tot; ++q) tmp[q] = 0.5 * (rhoIn[q] + rhoNew[q]);
  orc_energies(p, tmp, dt, C, L, u1, u2, tn, e);
  double Eh = orc_total_energy(p, e);
  double c1 = 4.0 * Eh - E1 - 3.0 * c0, c2 = E1 - c0 - c1, step;
  if (c2 > 0 && fabs(-c1 / 2.0 / c2 - 0.5) < 0.5) step = -c1 / 2.0 / c2;
  else if (E1 < c0) step = 1.0;
  else step = p->step_alt;
  for (i64 a = 0; a < n; ++a) dt[a] = (1 - step) * DensIn[a] + step * Dens[a] - 0.5;
  for (size_t q = 0; q < tot; ++q) tmp[q] = (1.0 - step) * rhoIn[q] + step * rhoNew[q];
  orc_energies(p, tmp, dt, C, L, u1, u2, tn, e_inout);
  free(dt); free(tmp);
  return step;
}

/* ------------------------------------------------------------------ SCF driver */

/* The SCF sequence of Main.f90:236-679 for numS = 1.
 *   rho_start   NULL: first diagonalisation (Main.f90:238-282); else the
 *               initial rho (as read from a checkpoint, Main.f90:187-235)
 *   occ_mode    0 reference T=0 rule with the zdeg quirk, 1 physical T=0,
 *               2 Fermi-Dirac at temperature kT (eV)
 *   mix_mode    0 ODA (OptimalStep), 1 linear with weight beta
 *   etol        >0: stop after the first iteration with |E_in - E_new| < etol
 *   traj        itmax x 12 row-major: it, E_in, 2Ekin_in, 2EF_in, EH_in, EU_in,
 *               maxdDensity, Step, E_in - E_new, E_Fermi, nOcc, nPart
 *   rot, c3perm NULL: Solve / GetFock; else nenforceC3 = 1 (Main.f90:244-250,
 *               273-276, 375-378, 500-505): Solve_C3 / GetFock_C3 with the
 *               layer rotation signs and the orc_c3_pairs permutation
 *   sym, c2perm nenforceT / C2 / C2T bits (orc_symmetrize) applied to each
 *               iteration's new density (Main.f90:514-534), C2 partners;
 *               bit 8: nenforceValley (orc_valley_transform with the layer
 *               rotation signs rot_v; reference weights)
 * Returns the number of iterations run (negative on an eigensolver error or,
 * with c3perm, numk % 3 != 0). */
i64 orc_scf(const orc_params* p, const double* C, const double* tn, const double* MomVal,
            const i64* nMomComp, const i64* u1, const i64* u2, const double* L,
            const zc* rho_start, i64 itmax, int occ_mode, double kT, int mix_mode, double beta,
            double etol, double* traj, zc* rho_out, double* timing, int nthreads,
            double* bands_last, const i64* rot, const i64* c3perm, int sym, const i64* c2perm,
            const i64* rot_v) {
  i64 n = p->ndim, nnc = p->nnc, Nk = p->numk * p->numk;
  i64 nb = p->nUpper - p->nLower + 1, N = nb * Nk;
  size_t nn = (size_t)n * n, tot = nn * nnc;
  zc* bulk = (zc*)malloc(sizeof(zc) * tot);
  zc* rho = (zc*)malloc(sizeof(zc) * tot);
  zc* rhoIn = (zc*)malloc(sizeof(zc) * tot);
  zc* zEig = (zc*)malloc(sizeof(zc) * n * N);
  zc* zSorted = (zc*)malloc(sizeof(zc) * n * N);
  i64* nSM = (i64*)malloc(sizeof(i64) * 2 * N);
  double* Bands = (double*)malloc(sizeof(double) * N);
  double* S = (double*)malloc(sizeof(double) * N);
  double* wts = (double*)malloc(sizeof(double) * N);
  double* V = (double*)calloc(n, sizeof(double));
  double* D = (double*)malloc(sizeof(double) * n);
  double* Din = (double*)malloc(sizeof(double) * n);
  double* dt = (double*)malloc(sizeof(double) * n);
  double e[4], ein[4], Etot, EtotIn, EF = 0, Deg = 0;
  i64 nF = 0, nOcc = 0, nPart = 0, it = 0;
  double t0 = omp_get_wtime(), tsolve = 0;
  int info;

#define GETFOCK(mode, w, cnt)                                                            \
  do {                                                                                   \
    if (c3perm)                                                                          \
      orc_getfock_c3(p, rot, c3perm, rho, zSorted, nSM, Deg, nOcc, nPart, C, MomVal, u1, u2, \
                     mode, w, cnt);                                                      \
    else                                                                                 \
      orc_getfock(p, rho, zSorted, nSM, Deg, nOcc, nPart, C, MomVal, u1, u2, mode, w, cnt); \
  } while (0)
#define SOLVE(rho_in)                                                                    \
  (c3perm ? orc_solve_c3(p, rot, c3perm, bulk, rho_in, V, Bands, zEig, C, MomVal, nMomComp, \
                         u1, u2, tn, nthreads)                                           \
          : orc_solve(p, bulk, rho_in, V, Bands, zEig, C, MomVal, nMomComp, u1, u2, tn, nthreads))
#define WINDOW_DENSITY()                                                             \
  do {                                                                               \
    orc_sort_energies(p, Bands, zEig, nMomComp, S, zSorted, nSM, NULL);              \
    orc_fermi(p, S, N, &nF, &EF, &nOcc, &nPart, &Deg);                               \
    memset(rho, 0, sizeof(zc) * tot);                                                \
    if (occ_mode == 2) {                                                             \
      EF = orc_fermi_dirac(S, N, nF, kT, wts);                                       \
      GETFOCK(2, wts, N);                                                            \
    } else {                                                                         \
      GETFOCK(occ_mode, NULL, 0);                                                    \
    }                                                                                \
    for (size_t q = 0; q < tot; ++q) rho[q] = (rho[q] + bulk[q]) / (zc)(double)Nk;  \
  } while (0)

  if (rho_start == NULL) {
    memset(rhoIn, 0, sizeof(zc) * tot);
    double t = omp_get_wtime();
    info = SOLVE(rhoIn);
    tsolve += omp_get_wtime() - t;
    if (info) return -1;
    WINDOW_DENSITY();
  } else {
    memcpy(rho, rho_start, sizeof(zc) * tot);
  }
  for (i64 a = 0; a < n; ++a) { D[a] = creal(rho[a + n * a]); dt[a] = D[a] - 0.5; }
  orc_energies(p, rho, dt, C, L, u1, u2, tn, e);
  Etot = orc_total_energy(p, e);
  orc_potential(p, L, D, V);
  memcpy(Din, D, sizeof(double) * n);
  memcpy(rhoIn, rho, sizeof(zc) * tot);

  while (it < itmax) {
    ++it;
    double t = omp_get_wtime();
    info = SOLVE(rho);
    tsolve += omp_get_wtime() - t;
    if (info) return -it;
    if (sym & 8) {
      /* nenforceValley (Main.f90:469-494): GetFock, ValleyTransform, sort,
       * GetFock again, rho = (rho / 2 + bulk) / Nk */
      orc_sort_energies(p, Bands, zEig, nMomComp, S, zSorted, nSM, NULL);
      orc_fermi(p, S, N, &nF, &EF, &nOcc, &nPart, &Deg);
      memset(rho, 0, sizeof(zc) * tot);
      GETFOCK(occ_mode, NULL, 0);
      orc_valley_transform(p, rot_v, C, tn, MomVal, nMomComp, u1, u2, zEig);
      orc_sort_energies(p, Bands, zEig, nMomComp, S, zSorted, nSM, NULL);
      GETFOCK(occ_mode, NULL, 0);
      for (size_t q = 0; q < tot; ++q) rho[q] = (0.5 * rho[q] + bulk[q]) / (zc)(double)Nk;
    } else {
      WINDOW_DENSITY();
    }
    if (sym & 7) orc_symmetrize(p, rho, sym & 7, c2perm);
    memcpy(ein, e, sizeof(e));
    EtotIn = Etot;
    for (i64 a = 0; a < n; ++a) D[a] = creal(rho[a + n * a]);
    double step;
    if (mix_mode == 0) {
      step = orc_optimal_step(p, e, rho, rhoIn, D, Din, C, L, u1, u2, tn);
    } else {
      step = beta;
      for (i64 a = 0; a < n; ++a) dt[a] = (1 - step) * Din[a] + step * D[a] - 0.5;
      for (size_t q = 0; q < tot; ++q) rho[q] = step * rho[q] + (1.0 - step) * rhoIn[q];
      orc_energies(p, rho, dt, C, L, u1, u2, tn, e);
    }
    Etot = orc_total_energy(p, e);
    if (mix_mode == 0)
      for (size_t q = 0; q < tot; ++q) rho[q] = step * rho[q] + (1.0 - step) * rhoIn[q];
    double dconv = 0;
    for (i64 a = 0; a < n; ++a) {
      D[a] = creal(rho[a + n * a]);
      if (dconv < fabs(D[a] - Din[a])) dconv = fabs(D[a] - Din[a]);
    }
    memcpy(Din, D, sizeof(double) * n);
    memcpy(rhoIn, rho, sizeof(zc) * tot);
    orc_potential(p, L, D, V);
    if (traj) {
      double* r = traj + 12 * (it - 1);
      r[0] = (double)it; r[1] = EtotIn; r[2] = 2 * ein[0]; r[3] = 2 * ein[1]; r[4] = ein[2]; r[5] = ein[3];
      r[6] = dconv; r[7] = step; r[8] = EtotIn - Etot; r[9] = EF; r[10] = (double)nOcc; r[11] = (double)nPart;
    }
    if (etol > 0 && fabs(EtotIn - Etot) < etol) break;
  }
#undef WINDOW_DENSITY
#undef SOLVE
#undef GETFOCK
  if (rho_out) memcpy(rho_out, rho, sizeof(zc) * tot);
  /* window bands of the last iteration's Solve, Bands(numb, Nk) as in Main.f90 */
  if (bands_last) memcpy(bands_last, Bands, sizeof(double) * N);
  if (timing) { timing[0] = omp_get_wtime() - t0; timing[1] = tsolve; }
  free(bulk); free(rho); free(rhoIn); free(zEig); free(zSorted); free(nSM); free(Bands);
  free(S); free(wts); free(V); free(D); free(Din); free(dt);
  return it;
}

/* ------------------------------------------------------------------ two spins (numS = 2) */

/* SortEnergies_2Spins, HartreeFock.f90:929-1004, with its window
 * W = [[NP-1, NP-1], [NP+3, NP+3]] (Main.f90:418-421, NP = NeutralityPoint).
 * S1: numb*Nk x 2 (the two spins' sorted energies); out: 2*numb*Nk x 2
 * (energy, spin label 1 or 2).  The reference fills its spin-2 lead-in from
 * row ncount1 of spin 2 while advancing ncount2 (HartreeFock.f90:951-956);
 * that is kept.  Its merge conditions read one row past the end of a column
 * before testing the row bound; the bound decides either way, so it is tested
 * first here. */
void orc_sort_2spins(const orc_params* p, const double* S1, double* out) {
  i64 Nk = p->numk * p->numk, N = (p->nUpper - p->nLower + 1) * Nk;
  i64 NP = (p->ndim / 2 - p->nLower) + 1;
  i64 w1 = NP - 1, w2 = NP + 3;
  i64 nb2 = 2 * N, cb = 0, c1 = 1, c2 = 1;
#define E1(c) S1[(c) - 1]
#define E2(c) S1[(c) - 1 + N]
#define PUT(v, lab) do { out[cb] = (v); out[cb + nb2] = (lab); ++cb; } while (0)
  for (i64 q = 1; q <= (w1 - 1) * Nk; ++q) { PUT(E1(c1), 1.0); ++c1; }
  for (i64 q = 1; q <= (w1 - 1) * Nk; ++q) { PUT(E2(c1), 2.0); ++c2; }
  while (c1 <= Nk * (w2 - 1) && c2 <= Nk * (w2 - 1)) {
    if (E1(c1) <= E2(c2)) {
      PUT(E1(c1), 1.0); ++c1;
      while (c1 <= N && E1(c1) <= E2(c2)) { PUT(E1(c1), 1.0); ++c1; }
      PUT(E2(c2), 2.0); ++c2;
    } else {
      PUT(E2(c2), 2.0); ++c2;
      while (c2 <= N && E2(c2) <= E1(c1)) { PUT(E2(c2), 2.0); ++c2; }
      PUT(E1(c1), 1.0); ++c1;
    }
  }
  while (c1 <= N) { PUT(E1(c1), 1.0); ++c1; }
  while (c2 <= N) { PUT(E2(c2), 2.0); ++c2; }
#undef E1
#undef E2
#undef PUT
}

/* The two-spin Fermi block, Main.f90:416-458: nF = 2 nint(Nk NP + Nk/2 nfilling)
 * over the merged list, per-spin occupied and partially occupied counts, and
 * one DegFactor (0 without partially occupied states). */
void orc_fermi_2spins(const orc_params* p, const double* S2, i64* nF, double* EF, i64* nOcc,
                      i64* nPart, double* Deg) {
  i64 Nk = p->numk * p->numk, N2 = 2 * (p->nUpper - p->nLower + 1) * Nk;
  i64 NP = (p->ndim / 2 - p->nLower) + 1;
  *nF = 2 * f_nint((double)(Nk * NP) + (double)Nk / 2.0 * p->nfilling);
  *EF = (S2[*nF - 1] + S2[*nF]) / 2.0;
  nOcc[0] = nOcc[1] = nPart[0] = nPart[1] = 0;
  for (i64 q = 0; q < N2; ++q) {
    int s = S2[q + N2] == 1.0 ? 0 : 1;
    if ((S2[q] - *EF) < -p->energy_tol) ++nOcc[s];
    else if (fabs(S2[q] - *EF) < p->energy_tol) { ++nPart[s]; ++nOcc[s]; }
  }
  if (nPart[0] != 0 || nPart[1] != 0)
    *Deg = 1.0 / (double)(nPart[0] + nPart[1]) *
           (double)(*nF + nPart[0] + nPart[1] - nOcc[0] - nOcc[1]);
  else
    *Deg = 0.0;
}

/* Energies of two spin densities (Main.f90:301-312, HartreeFock.f90:1162-1221
 * for numS = 2): e = {Ekin1 + Ekin2, EF1 + EF2, EH(d1, d2), EU(d1, d2)}; the
 * total is alpha EF + Ekin + U EU + alphaH EH (factor 3 - numS = 1). */
static void energies_2spins(const orc_params* p, const zc* z1, const zc* z2, const double* d1,
                            const double* d2, const double* C, const double* L, const i64* u1,
                            const i64* u2, const double* tn, double* e) {
  e[0] = orc_kinetic_energy(p, z1, C, u1, u2, tn) + orc_kinetic_energy(p, z2, C, u1, u2, tn);
  e[1] = orc_fock_energy(p, z1, C, u1, u2, tn) + orc_fock_energy(p, z2, C, u1, u2, tn);
  orc_hartree_hubbard(p, L, d1, d2, &e[2], &e[3]);
}

static double total_2spins(const orc_params* p, const double* e) {
  return p->alpha * e[1] + e[0] + p->U * e[3] + p->alphaH * e[2];
}

/* The SCF sequence of Main.f90:236-679 for numS = 2 (reference weights and
 * ODA mixing).  rho is n x n x nnc x 2.  rho_start NULL: the first
 * diagonalisation of spin 1 (Main.f90:238-288) copied to spin 2; else both
 * spins' densities (the reference's two checkpoint files, Main.f90:201-235).
 * traj rows as orc_scf: it, E_in, Ekin1+Ekin2, EF1+EF2, EH, EU (input state),
 * max over spins of max|dDensity|, Step, E_in - E_new, E_Fermi,
 * nOcc1 + nOcc2, nPart1 + nPart2 (Main.f90:616-639).  rot/c3perm and
 * sym/c2perm as in orc_scf (Solve_C3 / GetFock_C3 per spin; symmetrisation
 * of each spin's new density).  Returns the iterations run, negative on an
 * eigensolver error. */
i64 orc_scf2(const orc_params* p, const double* C, const double* tn, const double* MomVal,
             const i64* nMomComp, const i64* u1, const i64* u2, const double* L,
             const zc* rho_start, i64 itmax, double* traj, zc* rho_out, int nthreads,
             const i64* rot, const i64* c3perm, int sym, const i64* c2perm) {
  i64 n = p->ndim, nnc = p->nnc, Nk = p->numk * p->numk;
  i64 nb = p->nUpper - p->nLower + 1, N = nb * Nk;
  size_t nn = (size_t)n * n, tot = nn * nnc;
  zc* bulk = (zc*)malloc(sizeof(zc) * tot * 2);
  zc* rho = (zc*)malloc(sizeof(zc) * tot * 2);
  zc* rhoIn = (zc*)malloc(sizeof(zc) * tot * 2);
  zc* tmp = (zc*)malloc(sizeof(zc) * tot * 2);
  zc* zEig = (zc*)malloc(sizeof(zc) * n * N * 2);
  zc* zSorted = (zc*)malloc(sizeof(zc) * n * N * 2);
  i64* nSM = (i64*)malloc(sizeof(i64) * 2 * N * 2);
  double* Bands = (double*)malloc(sizeof(double) * N * 2);
  double* S1 = (double*)malloc(sizeof(double) * N * 2);
  double* S2 = (double*)malloc(sizeof(double) * N * 4);
  double* V = (double*)calloc(2 * n, sizeof(double));
  double* D = (double*)malloc(sizeof(double) * 2 * n);
  double* Din = (double*)malloc(sizeof(double) * 2 * n);
  double* dt = (double*)malloc(sizeof(double) * 2 * n);
  double e[4], ein[4], Etot, EtotIn, EF = 0, Deg = 0;
  i64 nF = 0, nOcc[2] = {0, 0}, nPart[2] = {0, 0}, it = 0;
  int info;
#define DENS(r, d) do { for (int s_ = 0; s_ < 2; ++s_) for (i64 a = 0; a < n; ++a) \
                        (d)[a + n * s_] = creal((r)[tot * s_ + a + n * a]); } while (0)
#define SUB(d, o) do { for (i64 a = 0; a < 2 * n; ++a) (o)[a] = (d)[a] - 0.5; } while (0)
#define ENERGIES(r, d, out) energies_2spins(p, (r), (r) + tot, (d), (d) + n, C, L, u1, u2, tn, (out))
#define POTENTIAL()                                                                     \
  do {                                                                                  \
    for (i64 a = 0; a < n; ++a) {                                                       \
      double s_ = 0;                                                                    \
      for (i64 b = 0; b < n; ++b)                                                       \
        s_ += (D[b] + D[b + n] - 2.0 * 0.5) * p->alphaH * L[a + n * b];                 \
      V[a + n] = s_ + p->U * (D[a] - 0.5);                                              \
      V[a] = s_ + p->U * (D[a + n] - 0.5);                                              \
    }                                                                                   \
  } while (0)

  if (rho_start == NULL) {
    memset(rhoIn, 0, sizeof(zc) * tot);
    info = c3perm ? orc_solve_c3(p, rot, c3perm, bulk, rhoIn, V, Bands, zEig, C, MomVal, nMomComp,
                                 u1, u2, tn, nthreads)
                  : orc_solve(p, bulk, rhoIn, V, Bands, zEig, C, MomVal, nMomComp, u1, u2, tn, nthreads);
    if (info) return -1;
    orc_sort_energies(p, Bands, zEig, nMomComp, S1, zSorted, nSM, NULL);
    i64 nF1, o1, q1;
    orc_fermi(p, S1, N, &nF1, &EF, &o1, &q1, &Deg);
    memset(rho, 0, sizeof(zc) * tot);
    if (c3perm)
      orc_getfock_c3(p, rot, c3perm, rho, zSorted, nSM, Deg, o1, q1, C, MomVal, u1, u2, 0, NULL, 0);
    else
      orc_getfock(p, rho, zSorted, nSM, Deg, o1, q1, C, MomVal, u1, u2, 0, NULL, 0);
    for (size_t q = 0; q < tot; ++q) rho[q] = (rho[q] + bulk[q]) / (zc)(double)Nk;
    memcpy(rho + tot, rho, sizeof(zc) * tot);
  } else {
    memcpy(rho, rho_start, sizeof(zc) * tot * 2);
  }
  DENS(rho, D);
  SUB(D, dt);
  ENERGIES(rho, dt, e);
  Etot = total_2spins(p, e);
  POTENTIAL();
  memcpy(Din, D, sizeof(double) * 2 * n);
  memcpy(rhoIn, rho, sizeof(zc) * tot * 2);

  while (it < itmax) {
    ++it;
    for (int s = 0; s < 2; ++s) {
      info = c3perm ? orc_solve_c3(p, rot, c3perm, bulk + tot * s, rho + tot * s, V + n * s,
                                   Bands + N * s, zEig + n * N * s, C, MomVal, nMomComp, u1, u2,
                                   tn, nthreads)
                    : orc_solve(p, bulk + tot * s, rho + tot * s, V + n * s, Bands + N * s,
                                zEig + n * N * s, C, MomVal, nMomComp, u1, u2, tn, nthreads);
      if (info) return -it;
      orc_sort_energies(p, Bands + N * s, zEig + n * N * s, nMomComp, S1 + N * s,
                        zSorted + n * N * s, nSM + 2 * N * s, NULL);
    }
    orc_sort_2spins(p, S1, S2);
    orc_fermi_2spins(p, S2, &nF, &EF, nOcc, nPart, &Deg);
    for (int s = 0; s < 2; ++s) {
      zc* r = rho + tot * s;
      memset(r, 0, sizeof(zc) * tot);
      if (c3perm)
        orc_getfock_c3(p, rot, c3perm, r, zSorted + n * N * s, nSM + 2 * N * s, Deg, nOcc[s],
                       nPart[s], C, MomVal, u1, u2, 0, NULL, 0);
      else
        orc_getfock(p, r, zSorted + n * N * s, nSM + 2 * N * s, Deg, nOcc[s], nPart[s], C, MomVal,
                    u1, u2, 0, NULL, 0);
      for (size_t q = 0; q < tot; ++q) r[q] = (r[q] + bulk[tot * s + q]) / (zc)(double)Nk;
      if (sym) orc_symmetrize(p, r, sym, c2perm);
    }
    memcpy(ein, e, sizeof(e));
    EtotIn = Etot;
    DENS(rho, D);
    /* OptimalStep for numS = 2 (HartreeFock.f90:1148-1227) */
    double c0 = EtotIn, t1, t2, step;
    SUB(D, dt);
    ENERGIES(rho, dt, e);
    t1 = total_2spins(p, e);
    for (i64 a = 0; a < 2 * n; ++a) dt[a] = 0.5 * (D[a] + Din[a]) - 0.5;
    for (size_t q = 0; q < 2 * tot; ++q) tmp[q] = 0.5 * (rhoIn[q] + rho[q]);
    ENERGIES(tmp, dt, e);
    t2 = total_2spins(p, e);
    double k1 = 4.0 * t2 - t1 - 3.0 * c0, k2 = t1 - c0 - k1;
    if (k2 > 0 && fabs(-k1 / 2.0 / k2 - 0.5) < 0.5) step = -k1 / 2.0 / k2;
    else if (t1 < c0) step = 1.0;
    else step = p->step_alt;
    for (i64 a = 0; a < 2 * n; ++a) dt[a] = (1 - step) * Din[a] + step * D[a] - 0.5;
    for (size_t q = 0; q < 2 * tot; ++q) tmp[q] = (1.0 - step) * rhoIn[q] + step * rho[q];
    ENERGIES(tmp, dt, e);
    Etot = total_2spins(p, e);
    for (size_t q = 0; q < 2 * tot; ++q) rho[q] = step * rho[q] + (1.0 - step) * rhoIn[q];
    DENS(rho, D);
    double dconv = 0;
    for (i64 a = 0; a < 2 * n; ++a)
      if (dconv < fabs(D[a] - Din[a])) dconv = fabs(D[a] - Din[a]);
    memcpy(Din, D, sizeof(double) * 2 * n);
    memcpy(rhoIn, rho, sizeof(zc) * tot * 2);
    POTENTIAL();
    if (traj) {
      double* r = traj + 12 * (it - 1);
      r[0] = (double)it; r[1] = EtotIn; r[2] = ein[0]; r[3] = ein[1]; r[4] = ein[2]; r[5] = ein[3];
      r[6] = dconv; r[7] = step; r[8] = EtotIn - Etot; r[9] = EF;
      r[10] = (double)(nOcc[0] + nOcc[1]); r[11] = (double)(nPart[0] + nPart[1]);
    }
  }
#undef DENS
#undef SUB
#undef ENERGIES
#undef POTENTIAL
  if (rho_out) memcpy(rho_out, rho, sizeof(zc) * tot * 2);
  free(bulk); free(rho); free(rhoIn); free(tmp); free(zEig); free(zSorted); free(nSM); free(Bands);
  free(S1); free(S2); free(V); free(D); free(Din); free(dt);
  return it;
}
