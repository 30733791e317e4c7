"""ctypes binding of the CPU restatement (oracle/forge_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libforge_oracle.so")

i64 = C.c_int64
dbl = C.c_double
P = C.c_void_p


class OrcParams(C.Structure):
    _fields_ = [(n, i64) for n in ("ndim", "nlayers", "numk", "numI", "numC", "nscreen",
                                   "fphase", "nnc", "nLower", "nUpper", "numb")] + \
               [(n, dbl) for n in ("nfilling", "U", "alpha", "alphaH", "Delta", "xi", "tz",
                                   "a0", "r0", "d0", "energy_tol", "step_alt")]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        # RTLD_GLOBAL: MKL dlopens its CPU-specific kernels, which need
        # mkl_core symbols visible globally.
        L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        L.orc_ft.restype = dbl
        L.orc_fv.restype = dbl
        L.orc_ndim.restype = i64
        L.orc_num_neighbor_cells.restype = i64
        L.orc_zdiagred.restype = C.c_int
        L.orc_solve.restype = C.c_int
        L.orc_solve_c3.restype = C.c_int
        L.orc_kinetic_energy.restype = dbl
        L.orc_fock_energy.restype = dbl
        L.orc_total_energy.restype = dbl
        L.orc_optimal_step.restype = dbl
        L.orc_fermi_dirac.restype = dbl
        L.orc_scf.restype = i64
        L.orc_scf2.restype = i64
        L.orc_c3_pairs.restype = i64
        L.orc_c2_pairs.restype = i64
        _lib = L
    return _lib


class Args:
    """Keeps every array converted for one foreign call alive until the call
    returns (a bare temporary could be freed before the C code reads it)."""

    def __init__(self):
        self.keep = []

    def __call__(self, a, dtype=None, fortran=True):
        if a is None:
            return None
        a = np.asarray(a, dtype=dtype)
        a = np.asfortranarray(a) if fortran else np.ascontiguousarray(a)
        self.keep.append(a)
        return C.c_void_p(a.ctypes.data)


def ptr(a):
    """Pointer to an array the caller holds (no conversion)."""
    if a is None:
        return None
    assert a.flags.f_contiguous or a.flags.c_contiguous
    return C.c_void_p(a.ctypes.data)


def fz(shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype, order="F")


class System:
    """Geometry + parameters of one reference configuration (Setup.f90)."""

    def __init__(self, ntheta, numk, nfilling=0.0, nlayers=2, xi_nm=40.0, U=4.0, epsilon=10.0,
                 Delta=0.0, numI=1, numC=3, nscreen=2, rotate=None):
        L = lib()
        self.ntheta, self.numk, self.nlayers = ntheta, numk, nlayers
        n = int(L.orc_ndim(i64(ntheta), i64(nlayers)))
        numb = min(20, n)
        ncb = numb // 2
        self.nnc = int(L.orc_num_neighbor_cells(i64(numI)))
        p = OrcParams()
        p.ndim, p.nlayers, p.numk, p.numI, p.numC, p.nscreen, p.fphase, p.nnc = \
            n, nlayers, numk, numI, numC, nscreen, 0, self.nnc
        p.nLower = (n - 2 * numb + 2 * ncb) // 2 + 1
        p.nUpper = (n + 2 * ncb) // 2
        p.numb = numb
        p.nfilling, p.U = nfilling, U
        p.alpha = p.alphaH = 5.853 / epsilon * 1.0
        p.Delta, p.xi, p.tz = Delta, xi_nm / 0.246, 1.35772
        p.a0, p.r0, p.d0 = 1.0 / np.sqrt(3.0), 0.184, 1.35772
        p.energy_tol, p.step_alt = 0.0001, 1.0
        self.p = p
        self.n = n
        if rotate is None:
            rotate = [-1, 1] if nlayers == 2 else ([-1, 1, -1] if nlayers == 3 else [-1])
        rot = np.array(rotate, dtype=np.int64)
        self.rot = rot
        Nk = numk * numk
        self.Coords = fz((n, 3))
        self.tn = fz((2, 3))
        self.g1 = np.zeros(2)
        self.g12 = np.zeros(2)
        self.MomentaValues = fz((numk, numk, 2))
        self.nMomentaComponents = fz((Nk, 2), np.int64)
        self.u1 = np.zeros(self.nnc, np.int64)
        self.u2 = np.zeros(self.nnc, np.int64)
        L.orc_geometry(i64(ntheta), i64(nlayers), ptr(rot), i64(numk), i64(numI), dbl(p.tz),
                       ptr(self.Coords), ptr(self.tn), ptr(self.g1), ptr(self.g12),
                       ptr(self.MomentaValues), ptr(self.nMomentaComponents), ptr(self.u1),
                       ptr(self.u2))
        self._L = None

    @property
    def pp(self):
        return C.byref(self.p)

    def vk(self, icount):
        i1, i2 = self.nMomentaComponents[icount]
        return np.array(self.MomentaValues[i1, i2, :])

    def ft(self, ri, rj):
        A = Args()
        return lib().orc_ft(self.pp, A(ri, float), A(rj, float))

    def fv(self, ri, rj):
        A = Args()
        return lib().orc_fv(self.pp, A(ri, float), A(rj, float))

    def longrange(self):
        if self._L is None:
            self._L = fz((self.n, self.n))
            lib().orc_longrange(self.pp, ptr(self.Coords), ptr(self.tn), ptr(self._L))
        return self._L

    def hamiltonian(self, vk, Potential=None, rho=None):
        zH = fz((self.n, self.n), np.complex128)
        A = Args()
        lib().orc_hamiltonian_hf(self.pp, ptr(zH), ptr(self.Coords), A(Potential, float),
                                 A(rho, np.complex128), ptr(self.u1), ptr(self.u2), A(vk, float),
                                 ptr(self.tn))
        return zH

    def zdiagred(self, H, il, iu):
        Hc = np.array(H, dtype=np.complex128, order="F", copy=True)
        w = np.zeros(self.n)
        info = lib().orc_zdiagred(ptr(Hc), i64(self.n), ptr(w), i64(il), i64(iu))
        return w[:iu - il + 1], Hc[:, :iu - il + 1], info

    def solve(self, rho, Potential, nthreads=0):
        p = self.p
        Nk = self.numk ** 2
        nb = p.nUpper - p.nLower + 1
        bulk = fz((self.n, self.n, self.nnc), np.complex128)
        Bands = fz((nb, Nk))
        zEig = fz((self.n, nb, Nk), np.complex128)
        A = Args()
        info = lib().orc_solve(self.pp, ptr(bulk), A(rho, np.complex128),
                               A(Potential, float), ptr(Bands), ptr(zEig),
                               ptr(self.Coords), ptr(self.MomentaValues),
                               ptr(self.nMomentaComponents), ptr(self.u1), ptr(self.u2),
                               ptr(self.tn), C.c_int(nthreads))
        return bulk, Bands, zEig, info

    def solve_c3(self, rho, Potential, nthreads=0):
        """Solve_C3: (bulk, Bands, zEig, info); info -2 when numk % 3 != 0."""
        perm, _ = self.c3_pairs()
        p = self.p
        Nk = self.numk ** 2
        nb = p.nUpper - p.nLower + 1
        bulk = fz((self.n, self.n, self.nnc), np.complex128)
        Bands = fz((nb, Nk))
        zEig = fz((self.n, nb, Nk), np.complex128)
        A = Args()
        info = lib().orc_solve_c3(self.pp, ptr(self.rot), ptr(perm), ptr(bulk),
                                  A(rho, np.complex128), A(Potential, float), ptr(Bands), ptr(zEig),
                                  ptr(self.Coords), ptr(self.MomentaValues),
                                  ptr(self.nMomentaComponents), ptr(self.u1), ptr(self.u2),
                                  ptr(self.tn), C.c_int(nthreads))
        return bulk, Bands, zEig, info

    def getfock_c3(self, zS, nSM, Deg, nOcc, nPart, occ_mode=0, weights=None):
        perm, _ = self.c3_pairs()
        rho = fz((self.n, self.n, self.nnc), np.complex128)
        cnt = 0 if weights is None else weights.size
        A = Args()
        lib().orc_getfock_c3(self.pp, ptr(self.rot), ptr(perm), ptr(rho), A(zS, np.complex128),
                             A(nSM, np.int64), dbl(Deg), i64(nOcc), i64(nPart), ptr(self.Coords),
                             ptr(self.MomentaValues), ptr(self.u1), ptr(self.u2),
                             C.c_int(occ_mode), A(weights, float), i64(cnt))
        return rho

    def valley_phase(self, vk):
        V = fz((self.n, self.n), np.complex128)
        A = Args()
        lib().orc_valley_phase(self.pp, ptr(self.rot), ptr(self.Coords), ptr(self.tn), ptr(self.u1),
                               ptr(self.u2), A(vk, float), ptr(V))
        return V

    def valley_transform(self, zEig):
        out = np.array(zEig, dtype=np.complex128, order="F", copy=True)
        lib().orc_valley_transform(self.pp, ptr(self.rot), ptr(self.Coords), ptr(self.tn),
                                   ptr(self.MomentaValues), ptr(self.nMomentaComponents),
                                   ptr(self.u1), ptr(self.u2), ptr(out))
        return out

    def sort(self, Bands, zEig):
        nb, Nk = Bands.shape
        N = nb * Nk
        S = np.zeros(N)
        zS = fz((self.n, N), np.complex128)
        nSM = fz((N, 2), np.int64)
        idx = np.zeros(N, np.int64)
        A = Args()
        lib().orc_sort_energies(self.pp, A(Bands, float), A(zEig, np.complex128),
                                ptr(self.nMomentaComponents), ptr(S), ptr(zS), ptr(nSM), ptr(idx))
        return S, zS, nSM, idx

    def fermi(self, S):
        nF, nOcc, nPart = i64(), i64(), i64()
        EF, Deg = dbl(), dbl()
        A = Args()
        lib().orc_fermi(self.pp, A(S, float), i64(np.size(S)), C.byref(nF), C.byref(EF), C.byref(nOcc),
                        C.byref(nPart), C.byref(Deg))
        return nF.value, EF.value, nOcc.value, nPart.value, Deg.value

    def getfock(self, zS, nSM, Deg, nOcc, nPart, occ_mode=0, weights=None):
        rho = fz((self.n, self.n, self.nnc), np.complex128)
        cnt = 0 if weights is None else weights.size
        A = Args()
        lib().orc_getfock(self.pp, ptr(rho), A(zS, np.complex128), A(nSM, np.int64), dbl(Deg),
                          i64(nOcc), i64(nPart), ptr(self.Coords), ptr(self.MomentaValues),
                          ptr(self.u1), ptr(self.u2), C.c_int(occ_mode), A(weights, float), i64(cnt))
        return rho

    def energies(self, rho, dens=None):
        if dens is None:
            dens = np.real(np.diag(rho[:, :, 0])) - 0.5
        out = np.zeros(4)
        A = Args()
        lib().orc_energies(self.pp, A(rho, np.complex128), A(dens, float), ptr(self.Coords),
                           ptr(self.longrange()), ptr(self.u1), ptr(self.u2), ptr(self.tn), ptr(out))
        return out

    def total_energy(self, e):
        A = Args()
        return lib().orc_total_energy(self.pp, A(e, float))

    def potential(self, Density):
        V = np.zeros(self.n)
        A = Args()
        lib().orc_potential(self.pp, ptr(self.longrange()), A(Density, float), ptr(V))
        return V

    def optimal_step(self, e_in, rhoNew, rhoIn, D, Din):
        e = np.array(e_in, dtype=float)
        A = Args()
        step = lib().orc_optimal_step(self.pp, ptr(e), A(rhoNew, np.complex128),
                                      A(rhoIn, np.complex128), A(D, float), A(Din, float),
                                      ptr(self.Coords),
                                      ptr(self.longrange()), ptr(self.u1), ptr(self.u2), ptr(self.tn))
        return step, e

    def scf2(self, itmax, rho_start=None, nthreads=0, c3=False, sym=0):
        """The numS = 2 SCF (reference weights, ODA): rho_start (n, n, nnc, 2)
        or None for the first diagonalisation copied to both spins; c3 and sym
        as in scf.  Returns (iterations, trajectory rows, rho (n, n, nnc, 2))."""
        perm = self.c3_pairs()[0] if c3 else None
        c2 = self.c2_pairs()[0] if sym else None
        traj = np.zeros((itmax, 12))
        rho_out = np.zeros((self.n, self.n, self.nnc, 2), np.complex128, order="F")
        A = Args()
        its = lib().orc_scf2(self.pp, ptr(self.Coords), ptr(self.tn), ptr(self.MomentaValues),
                             ptr(self.nMomentaComponents), ptr(self.u1), ptr(self.u2),
                             ptr(self.longrange()), A(rho_start, np.complex128), i64(itmax),
                             ptr(traj), ptr(rho_out), C.c_int(nthreads),
                             ptr(self.rot) if c3 else None, ptr(perm) if c3 else None,
                             C.c_int(sym), ptr(c2) if sym else None)
        return int(its), traj[:max(int(its), 0)], rho_out

    def sort_2spins(self, S1):
        """SortEnergies_2Spins of the two spins' sorted energies (numb*Nk x 2)."""
        N = S1.shape[0]
        out = np.zeros((2 * N, 2), order="F")
        A = Args()
        lib().orc_sort_2spins(self.pp, A(S1, float), ptr(out))
        return out

    def c3_pairs(self):
        """nC3pairs (C3_RelatedPoints, Geometry.f90:1172-1223), 0-based, and
        the number of sites matched."""
        perm = np.zeros(self.n, np.int64)
        cnt = int(lib().orc_c3_pairs(ptr(self.Coords), i64(self.n), i64(self.nlayers), ptr(perm)))
        return perm, cnt

    def symmetrize(self, rho, sym):
        """orc_symmetrize: Main.f90:514-534 on a copy of rho."""
        c2, _ = self.c2_pairs()
        out = np.array(rho, dtype=np.complex128, order="F", copy=True)
        lib().orc_symmetrize(self.pp, ptr(out), C.c_int(sym), ptr(c2))
        return out

    def c2_pairs(self):
        """nC2pairs (C2_RelatedPoints, Geometry.f90:1140-1167), 0-based, and
        the number of pairs matched."""
        perm = np.zeros(self.n, np.int64)
        cnt = int(lib().orc_c2_pairs(ptr(self.Coords), i64(self.n), i64(self.nlayers), ptr(perm)))
        return perm, cnt

    def scf(self, itmax, rho_start=None, occ_mode=0, kT=0.0, mix_mode=0, beta=0.3, etol=0.0,
            nthreads=0, keep_bands=False, c3=False, sym=0):
        """keep_bands: the last iteration's window bands (numb x Nk) are left
        in self.last_bands.  c3: the reference's nenforceC3 = 1 path
        (Solve_C3 / GetFock_C3).  sym: nenforceT (1) | nenforceC2 (2) |
        nenforceC2T (4) | nenforceValley (8)."""
        c2 = None
        if sym & 7:
            c2, cnt = self.c2_pairs()
            if cnt != self.n // 2:
                raise ValueError(f"cell is not C2-symmetric ({cnt} of {self.n // 2} pairs)")
        perm = None
        if c3:
            perm, cnt = self.c3_pairs()
            if cnt != self.n:
                raise ValueError(f"cell is not C3-symmetric ({cnt} of {self.n} sites matched)")
        traj = np.zeros((itmax, 12))
        rho_out = fz((self.n, self.n, self.nnc), np.complex128)
        timing = np.zeros(2)
        nb = int(self.p.nUpper - self.p.nLower + 1)
        bands = fz((nb, self.numk * self.numk), np.float64) if keep_bands else None
        A = Args()
        its = lib().orc_scf(self.pp, ptr(self.Coords), ptr(self.tn), ptr(self.MomentaValues),
                            ptr(self.nMomentaComponents), ptr(self.u1), ptr(self.u2),
                            ptr(self.longrange()), A(rho_start, np.complex128),
                            i64(itmax), C.c_int(occ_mode), dbl(kT), C.c_int(mix_mode), dbl(beta),
                            dbl(etol), ptr(traj), ptr(rho_out), ptr(timing), C.c_int(nthreads),
                            ptr(bands) if keep_bands else None,
                            ptr(self.rot) if c3 else None, ptr(perm) if c3 else None,
                            C.c_int(sym), ptr(c2) if sym & 7 else None, ptr(self.rot))
        if keep_bands:
            self.last_bands = bands
        return int(its), traj[:max(int(its), 0)], rho_out, timing
