"""Python host of the MI355X Hartree-Fock engine.

`Engine` mirrors the reference's hot-path subroutines (HartreeFock.f90,
TightBinding.f90, LapackRoutines.f90) with the same argument meaning, array
layout (Fortran order, complex128) and unnormalised outputs, on top of the
C ABI in include/forge_hf.h.  `scf_*` drive the device-resident loop of
Main.f90:236-679.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import lib as L
from .geometry import Geometry, build_geometry, ndim_of


@dataclass
class Setup:
    """Runtime form of Setup.f90's parameters."""
    ntheta: int = 5
    nlayers: int = 2
    numk: int = 6
    nfilling: float = 0.0
    U: float = 4.0
    epsilon: float = 10.0
    Delta: float = 0.0
    xi_nm: float = 40.0        # xi * 0.246: Setup.f90:22 expresses xi in units of a
    numI: int = 1
    numC: int = 3
    nscreen: int = 2
    nrelax: int = 0            # 1: relaxed bilayer lattice (Setup.f90:12, Main.f90:156-160)
    energy_tol: float = 1e-4   # EnergyTolerance, Setup.f90:51
    step_alt: float = 1.0      # StepAlternative, Setup.f90:52

    @property
    def ndim(self):
        return ndim_of(self.ntheta, self.nlayers)

    def params(self, nnc: int) -> L.ForgeParams:
        n = self.ndim
        numb = min(20, n)
        ncb = numb // 2
        p = L.ForgeParams()
        p.ndim, p.nlayers, p.numk, p.numI, p.numC = n, self.nlayers, self.numk, self.numI, self.numC
        p.nscreen, p.fphase, p.nnc = self.nscreen, 0, nnc
        p.nLower = (n - 2 * numb + 2 * ncb) // 2 + 1     # Setup.f90:67
        p.nUpper = (n + 2 * ncb) // 2                    # Setup.f90:68
        p.numb = numb
        p.nfilling, p.U = self.nfilling, self.U
        p.alpha = p.alphaH = 5.853 / self.epsilon * 1.0  # Setup.f90:58-59
        p.Delta, p.xi, p.tz = self.Delta, self.xi_nm / 0.246, 1.35772
        p.a0, p.r0, p.d0 = 1.0 / np.sqrt(3.0), 0.184, 1.35772
        p.energy_tol, p.step_alt = self.energy_tol, self.step_alt
        return p


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _f(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


class Engine:
    def __init__(self, setup: Setup, geom: Geometry | None = None):
        self.setup = setup
        self.geom = geom or build_geometry(setup.ntheta, setup.nlayers, setup.numk, setup.numI,
                                           nrelax=setup.nrelax)
        g = self.geom
        self.p = setup.params(g.nnc)
        self.n = g.ndim
        self.nnc = g.nnc
        self.Nk = setup.numk ** 2
        self.nb = self.p.nUpper - self.p.nLower + 1
        self._keep = [g.Coords, g.tn, g.MomentaValues, g.nMomentaComponents, g.nUnitCell_1,
                      g.nUnitCell_2]
        self.lib = L.load()
        ctx = C.c_void_p()
        L.check(self.lib.forge_ctx_create(C.byref(self.p), _p(g.Coords), _p(g.tn),
                                          _p(g.MomentaValues), _p(g.nMomentaComponents),
                                          _p(g.nUnitCell_1), _p(g.nUnitCell_2), C.byref(ctx)),
                "forge_ctx_create")
        self.ctx = ctx

    def close(self):
        if self.ctx:
            self.lib.forge_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference interfaces ------------------------------------------------
    def longrange(self):
        """LongRangeInteraction (HartreeFock.f90:14-53)."""
        out = np.zeros((self.n, self.n), order="F")
        L.check(self.lib.forge_longrange(self.ctx, _p(out)), "forge_longrange")
        return out

    def tb_hamiltonian(self, vk):
        """TightBindingHamiltonian (TightBinding.f90:13-55): lower triangle."""
        zH = np.zeros((self.n, self.n), np.complex128, order="F")
        v = _f(vk, np.float64)
        L.check(self.lib.forge_tb_hamiltonian(self.ctx, _p(zH), _p(v)), "forge_tb_hamiltonian")
        return zH

    def hamiltonian_hf(self, vk, Potential, zFock):
        """HamiltonianHartreeFock (HartreeFock.f90:57-127): lower triangle."""
        zH = np.zeros((self.n, self.n), np.complex128, order="F")
        v, pot, rho = _f(vk, np.float64), _f(Potential, np.float64), _f(zFock, np.complex128)
        L.check(self.lib.forge_hamiltonian_hf(self.ctx, _p(zH), _p(pot), _p(rho), _p(v)),
                "forge_hamiltonian_hf")
        return zH

    def solve(self, zFock, Potential):
        """Solve (HartreeFock.f90:402-483): (zFockBulk unnormalised, Bands, zEigenvectors)."""
        bulk = np.zeros((self.n, self.n, self.nnc), np.complex128, order="F")
        bands = np.zeros((self.nb, self.Nk), order="F")
        vecs = np.zeros((self.n, self.nb, self.Nk), np.complex128, order="F")
        rho, pot = _f(zFock, np.complex128), _f(Potential, np.float64)
        rc = L.check(self.lib.forge_solve(self.ctx, _p(bulk), _p(rho), _p(pot), _p(bands), _p(vecs)),
                     "forge_solve")
        if rc > 0:
            print("zheev failed", rc)  # the reference's reaction, LapackRoutines.f90:430
        return bulk, bands, vecs

    def getfock(self, zFock, zSortedEigenvectors, nSortedMomenta, DegFactor, OccStates,
                PartOccStates):
        """GetFock (HartreeFock.f90:487-560): returns zFock + window density."""
        out = np.array(zFock, dtype=np.complex128, order="F", copy=True)
        vecs = _f(zSortedEigenvectors, np.complex128)
        mom = _f(nSortedMomenta, np.int64)
        L.check(self.lib.forge_getfock(self.ctx, _p(out), _p(vecs), _p(mom), float(DegFactor),
                                       int(OccStates), int(PartOccStates)), "forge_getfock")
        return out

    # -- C3-reduced k-set (nenforceC3 = 1) -----------------------------------
    def c3_pairs(self):
        """nC3pairs of C3_RelatedPoints (Geometry.f90:1172-1223), 1-based."""
        g = self.geom
        out = np.zeros(self.n, np.int64)
        cnt = self.lib.forge_c3_pairs(_p(g.Coords), self.n, g.nlayers, _p(out))
        if cnt < 0:
            raise RuntimeError("forge_c3_pairs: " + L.last_error())
        if cnt != self.n:
            raise ValueError(f"cell is not C3-symmetric ({cnt} of {self.n} sites matched)")
        return out

    def _c3_tables(self):
        if getattr(self, "_c3", None) is None:
            rot = np.ascontiguousarray(self.geom.RotateLayers, dtype=np.int64)
            self._c3 = (rot, self.c3_pairs())
        return self._c3

    def solve_c3(self, zFock, Potential):
        """Solve_C3 (HartreeFock.f90:131-398): one diagonalisation per C3 orbit
        of momenta; (zFockBulk unnormalised, Bands, zEigenvectors) for all Nk."""
        rot, pairs = self._c3_tables()
        bulk = np.zeros((self.n, self.n, self.nnc), np.complex128, order="F")
        bands = np.zeros((self.nb, self.Nk), order="F")
        vecs = np.zeros((self.n, self.nb, self.Nk), np.complex128, order="F")
        rho, pot = _f(zFock, np.complex128), _f(Potential, np.float64)
        rc = L.check(self.lib.forge_solve_c3(self.ctx, _p(bulk), _p(rho), _p(pot), _p(bands), _p(vecs),
                                             _p(rot), _p(pairs)), "forge_solve_c3")
        if rc > 0:
            print("zheev failed", rc)
        return bulk, bands, vecs

    def getfock_c3(self, zFock, zSortedEigenvectors, nSortedMomenta, DegFactor, OccStates,
                   PartOccStates):
        """GetFock_C3 (HartreeFock.f90:564-881): returns zFock + window density,
        states at Gamma, K, K' symmetrised over their C3 images."""
        rot, pairs = self._c3_tables()
        out = np.array(zFock, dtype=np.complex128, order="F", copy=True)
        vecs = _f(zSortedEigenvectors, np.complex128)
        mom = _f(nSortedMomenta, np.int64)
        L.check(self.lib.forge_getfock_c3(self.ctx, _p(out), _p(vecs), _p(mom), float(DegFactor),
                                          int(OccStates), int(PartOccStates), _p(rot), _p(pairs)),
                "forge_getfock_c3")
        return out

    def set_c3(self, on=True):
        """nenforceC3 for the device-resident SCF loop (Main.f90:244-276, 375-505)."""
        if on:
            rot, pairs = self._c3_tables()
            L.check(self.lib.forge_scf_set_c3(self.ctx, _p(rot), _p(pairs)), "forge_scf_set_c3")
        else:
            L.check(self.lib.forge_scf_set_c3(self.ctx, None, None), "forge_scf_set_c3")

    # -- symmetry enforcement (nenforceT / nenforceC2 / nenforceC2T) ----------
    def c2_pairs(self):
        """nC2pairs of C2_RelatedPoints (Geometry.f90:1140-1167), 1-based."""
        g = self.geom
        out = np.zeros(self.n, np.int64)
        cnt = self.lib.forge_c2_pairs(_p(g.Coords), self.n, g.nlayers, _p(out))
        if cnt < 0:
            raise RuntimeError("forge_c2_pairs: " + L.last_error())
        if cnt != self.n // 2:
            raise ValueError(f"cell is not C2-symmetric ({cnt} of {self.n // 2} pairs)")
        return out

    def _c2_table(self):
        if getattr(self, "_c2", None) is None:
            self._c2 = self.c2_pairs()
        return self._c2

    def symmetrize(self, zFock, flags):
        """Main.f90:514-534 on a density: flags 1 (T) | 2 (C2) | 4 (C2T)."""
        out = np.array(zFock, dtype=np.complex128, order="F", copy=True)
        pairs = self._c2_table() if flags & 6 else None
        L.check(self.lib.forge_symmetrize(self.ctx, _p(out), int(flags), _p(pairs)), "forge_symmetrize")
        return out

    def set_symmetry(self, flags):
        """Symmetrisation of each new density in the device-resident loop."""
        pairs = self._c2_table() if flags & 6 else None
        L.check(self.lib.forge_scf_set_symmetry(self.ctx, int(flags), _p(pairs)),
                "forge_scf_set_symmetry")

    # -- valley projection (nenforceValley) --------------------------------------
    def _rot(self):
        return np.ascontiguousarray(self.geom.RotateLayers, dtype=np.int64)

    def valley_phase(self, vk):
        """ValleyPhase (TightBinding.f90:60-165): the valley operator at vk."""
        out = np.zeros((self.n, self.n), np.complex128, order="F")
        v, rot = _f(vk, np.float64), self._rot()
        L.check(self.lib.forge_valley_phase(self.ctx, _p(v), _p(rot), _p(out)), "forge_valley_phase")
        return out

    def valley_transform(self, zEigenvectors):
        """ValleyTransform (TightBinding.f90:170-211) of (n, numb, Nk) eigenvectors."""
        out = np.array(zEigenvectors, dtype=np.complex128, order="F", copy=True)
        rot = self._rot()
        L.check(self.lib.forge_valley_transform(self.ctx, _p(out), _p(rot)), "forge_valley_transform")
        return out

    def set_valley(self, on=True):
        """nenforceValley for the device-resident loop (Main.f90:469-494)."""
        rot = self._rot() if on else None
        L.check(self.lib.forge_scf_set_valley(self.ctx, _p(rot)), "forge_scf_set_valley")

    def energies(self, zFock, dens):
        """(Ekin, EF, EH, EU) of GetKineticEnergy/GetFockEnergy/GetHartreeHubbardEnergy."""
        out = np.zeros(4)
        rho, d = _f(zFock, np.complex128), _f(dens, np.float64)
        L.check(self.lib.forge_energies(self.ctx, _p(rho), _p(d), _p(out)), "forge_energies")
        return out

    def hartree_hubbard(self, dens1, dens2):
        """(fsumH, fsumU) of GetHartreeHubbardEnergy (HartreeFock.f90:1066-1092)."""
        out = np.zeros(2)
        d1, d2 = _f(dens1, np.float64), _f(dens2, np.float64)
        L.check(self.lib.forge_hartree_hubbard(self.ctx, _p(d1), _p(d2), _p(out)),
                "forge_hartree_hubbard")
        return out

    def total_energy(self, e):
        p = self.p
        return p.alpha * e[1] * 2.0 + e[0] * 2.0 + p.U * e[3] + p.alphaH * e[2]

    # -- device-resident SCF -----------------------------------------------
    def set_options(self, occ_mode=0, mix_mode=0, kT=0.0, beta=0.3, etol=0.0, precision=0,
                    timing=0, switch_de=0.0, graph=0, reduction=0):
        o = L.ScfOptions(occ_mode, mix_mode, kT, beta, etol, precision, timing, switch_de, graph,
                         reduction)
        L.check(self.lib.forge_scf_set_options(self.ctx, C.byref(o)), "forge_scf_set_options")

    def scf_init(self, rho_start=None):
        rho = None if rho_start is None else _f(rho_start, np.complex128)
        return L.check(self.lib.forge_scf_init(self.ctx, _p(rho)), "forge_scf_init")

    def scf_init_guess(self, rho_guess):
        """Start from the occupied density of H[rho_guess] (one diagonalisation)."""
        rho = _f(rho_guess, np.complex128)
        return L.check(self.lib.forge_scf_init_guess(self.ctx, _p(rho)), "forge_scf_init_guess")

    def scf_init_update(self, diag, row, col, cell, delta):
        """The current state with the guess applied (diagonal values and
        distinct (row, col, cell) increments, 0-based) becomes the start state,
        without a diagonalisation."""
        d = _f(diag, np.float64)
        r, c, k = (np.ascontiguousarray(a, np.int64) for a in (row, col, cell))
        v = np.ascontiguousarray(delta, np.complex128)
        return L.check(self.lib.forge_scf_init_update(self.ctx, _p(d), int(r.size), _p(r), _p(c),
                                                      _p(k), _p(v)), "forge_scf_init_update")

    def scf_init_guess_update(self, diag, row, col, cell, delta):
        """As scf_init_guess, the guess given as changes to the current state
        (diagonal values and distinct (row, col, cell) increments, 0-based)."""
        d = _f(diag, np.float64)
        r, c, k = (np.ascontiguousarray(a, np.int64) for a in (row, col, cell))
        v = np.ascontiguousarray(delta, np.complex128)
        return L.check(self.lib.forge_scf_init_guess_update(self.ctx, _p(d), int(r.size), _p(r),
                                                            _p(c), _p(k), _p(v)),
                       "forge_scf_init_guess_update")

    def scf_run(self, max_iter):
        traj = np.zeros((max(max_iter, 1), 12))
        done = C.c_int64(0)
        rc = L.check(self.lib.forge_scf_run(self.ctx, int(max_iter), _p(traj), C.byref(done)),
                     "forge_scf_run")
        if rc > 0:
            print("zheev failed", rc)
        return traj[:done.value]

    def write_reference_outputs(self, outdir, traj, statename="SSS", append=False):
        """The reference's output/Energy* and output/Convergence* files
        (Main.f90:352-358, 616-631) for trajectory rows of scf_run."""
        from .checkpoint import run_parameters
        from .outputs import write_outputs
        st = self.setup
        params = run_parameters(statename, st.nfilling, 1, st.ntheta, st.nlayers, st.nrelax, st.epsilon,
                                st.U, st.nscreen, st.xi_nm, st.Delta, st.numI, st.numC, st.numk)
        return write_outputs(outdir, params, traj, append=append)

    def density(self):
        rho = np.zeros((self.n, self.n, self.nnc), np.complex128, order="F")
        L.check(self.lib.forge_scf_get_density(self.ctx, _p(rho)), "forge_scf_get_density")
        return rho

    def timings(self):
        out = np.zeros(24)
        L.check(self.lib.forge_scf_timings(self.ctx, _p(out)), "forge_scf_timings")
        return out


class SpinPair:
    """The numS = 2 SCF on the device (forge_scf2_*): two engines of the same
    system, spin 1 and spin 2, owned by the caller.  Set the same options on
    both (occ_mode 0 or 1, mix_mode 0) before init()."""

    def __init__(self, spin1: Engine, spin2: Engine):
        self.engines = (spin1, spin2)
        self.lib = L.load()
        pr = C.c_void_p()
        L.check(self.lib.forge_scf2_create(spin1.ctx, spin2.ctx, C.byref(pr)), "forge_scf2_create")
        self.pr = pr

    def init(self, rho_start=None):
        """rho_start: both spins' densities as a sequence of two (n, n, nnc)
        arrays or one (n, n, nnc, 2) array; None: the first diagonalisation."""
        if rho_start is None:
            return L.check(self.lib.forge_scf2_init(self.pr, None), "forge_scf2_init")
        if isinstance(rho_start, (list, tuple)):
            rho_start = np.stack([np.asarray(r) for r in rho_start], axis=-1)
        rho = _f(rho_start, np.complex128)
        return L.check(self.lib.forge_scf2_init(self.pr, _p(rho)), "forge_scf2_init")

    def run(self, max_iter):
        """-> trajectory rows (it, E_in, Ekin1+Ekin2, EF1+EF2, EH, EU, max|dDensity|,
        Step, E_in - E_new, E_Fermi, nOcc, nPart)."""
        traj = np.zeros((max(max_iter, 1), 12))
        done = C.c_int64(0)
        rc = L.check(self.lib.forge_scf2_run(self.pr, int(max_iter), _p(traj), C.byref(done)),
                     "forge_scf2_run")
        if rc > 0:
            print("zheev failed", rc)
        return traj[:done.value]

    def density(self):
        e = self.engines[0]
        rho = np.zeros((e.n, e.n, e.nnc, 2), np.complex128, order="F")
        L.check(self.lib.forge_scf2_get_density(self.pr, _p(rho)), "forge_scf2_get_density")
        return rho

    def close(self):
        if self.pr:
            self.lib.forge_scf2_destroy(self.pr)
            self.pr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Sweep:
    """SCF of several systems of different orders advanced together, their
    eigensolves batched into one uneven-size reduction per iteration
    (forge_sweep_*).  The engines stay owned by the caller; set their options
    (FP64) and initial states before run()."""

    def __init__(self, engines):
        self.engines = list(engines)
        self.lib = L.load()
        arr = (C.c_void_p * len(self.engines))(*[e.ctx.value for e in self.engines])
        sw = C.c_void_p()
        L.check(self.lib.forge_sweep_create(arr, len(self.engines), C.byref(sw)),
                "forge_sweep_create")
        self.sw = sw

    def run(self, max_iter):
        """-> list of per-system trajectories (rows as Engine.scf_run)."""
        ns = len(self.engines)
        traj = np.zeros((ns, max(max_iter, 1), 12))
        done = np.zeros(ns, np.int64)
        rc = L.check(self.lib.forge_sweep_run(self.sw, int(max_iter), _p(traj), _p(done)),
                     "forge_sweep_run")
        if rc > 0:
            print("zheev failed", rc)
        return [traj[i, :done[i]] for i in range(ns)]

    def close(self):
        if self.sw:
            self.lib.forge_sweep_destroy(self.sw)
            self.sw = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def zdiagred(A, il, iu):
    """zdiagRed (LapackRoutines.f90:381-433) on the GPU: eigenpairs il..iu of the
    Hermitian matrix whose lower triangle A holds -> (eigenvalues, vectors, info)."""
    lib = L.load()
    n = A.shape[0]
    a = np.array(A, dtype=np.complex128, order="F", copy=True)
    w = np.zeros(n)
    info = C.c_int32(0)
    L.check(lib.forge_zdiagred(_p(a), n, _p(w), int(il), int(iu), C.byref(info)), "forge_zdiagred")
    m = iu - il + 1
    return w[:m], a[:, :m], info.value


def zdiagred_noalloc(A, il, iu):
    """zdiagred without device allocation (forge_zdiagred_noalloc, n <= 512)."""
    lib = L.load()
    n = A.shape[0]
    a = np.array(A, dtype=np.complex128, order="F", copy=True)
    w = np.zeros(n)
    info = C.c_int32(0)
    L.check(lib.forge_zdiagred_noalloc(_p(a), n, _p(w), int(il), int(iu), C.byref(info)),
            "forge_zdiagred_noalloc")
    m = iu - il + 1
    return w[:m], a[:, :m], info.value


def z_vector_mul(A, C, z_factor=None):
    """z_vector_mul(A, C, z_factor) (LapackRoutines.f90:168-183): C + z A A^H,
    A n x k (the reference passes one column)."""
    lib = L.load()
    A = _f(np.atleast_2d(np.asarray(A).T).T, np.complex128)
    out = _f(C, np.complex128).copy(order="F")
    n, k = A.shape
    z = None if z_factor is None else np.array([z_factor], np.complex128)
    L.check(lib.forge_z_vector_mul(_p(out), _p(A), n, k, _p(z)), "forge_z_vector_mul")
    return out


def z_matrix_mul(A, B):
    """z_matrix_mul(A, B, C) (LapackRoutines.f90:208-220): A B."""
    lib = L.load()
    A, B = _f(A, np.complex128), _f(B, np.complex128)
    m, k = A.shape
    n = B.shape[1]
    out = np.zeros((m, n), np.complex128, order="F")
    L.check(lib.forge_z_matrix_mul(_p(out), _p(A), _p(B), m, k, n), "forge_z_matrix_mul")
    return out


def z_matrix_overlap(A, B, z_factor=None):
    """z_matrix_overlap(A, B, C, z_factor) (LapackRoutines.f90:119-139): z A^H B."""
    lib = L.load()
    A, B = _f(A, np.complex128), _f(B, np.complex128)
    m1, m0 = A.shape
    out = np.zeros((m0, m0), np.complex128, order="F")
    z = None if z_factor is None else np.array([z_factor], np.complex128)
    L.check(lib.forge_z_matrix_overlap(_p(out), _p(A), _p(B), m1, m0, _p(z)),
            "forge_z_matrix_overlap")
    return out


def c_matrix_mul(A, B):
    """c_matrix_mul(A, B, C) (LapackRoutines.f90:224-236): complex(4) A B."""
    lib = L.load()
    A, B = _f(A, np.complex64), _f(B, np.complex64)
    m, k = A.shape
    n = B.shape[1]
    out = np.zeros((m, n), np.complex64, order="F")
    L.check(lib.forge_c_matrix_mul(_p(out), _p(A), _p(B), m, k, n), "forge_c_matrix_mul")
    return out
