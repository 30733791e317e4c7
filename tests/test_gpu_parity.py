"""GPU parity: every entry point of the C ABI against the reference fixtures
and the CPU oracle on the same inputs (FP64 path).

Tolerances: assembly and pair tables are elementwise FP64 (1e-12 absolute on
O(1-10) eV entries); eigenvalues 1e-10 eV (the reference metric asks 1e-8);
density matrices are compared as gauge-invariant sums (1e-9); SCF energies to
the reference's printed 1e-9 eV resolution plus accumulated rounding (1e-8).
"""
import sys

import numpy as np
import pytest

from golden_io import case_params, load_case, load_traj
from forge_amd import lib as L
from forge_amd.hf import Engine, Setup, zdiagred
from forge_amd.seed import seeded_rho

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_eng = {}


def engine_for(case):
    if case not in _eng:
        c = case_params(load_case(case))
        s = Setup(ntheta=c["ntheta"], nlayers=c["nlayers"], numk=c["numk"], nfilling=c["nfilling"],
                  xi_nm=c["xi_nm"])
        _eng[case] = Engine(s)
    return _eng[case]


def full(H):
    return np.tril(H) + np.tril(H, -1).conj().T


def test_device_present():
    assert L.device_count() >= 1


@pytest.mark.parametrize("case", ["t1k3", "t2k3", "t3k6", "tri1k3"])
def test_longrange_matches_reference(case):
    say("longrange", case)
    e = engine_for(case)
    np.testing.assert_allclose(e.longrange(), load_case(case)["LongRange"], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("case", ["t1k3", "t2k3", "t2k3xi10", "tri1k3"])
def test_hamiltonians_match_reference(case):
    say("hamiltonian", case)
    f = load_case(case)
    e = engine_for(case)
    Nk = e.Nk
    for tag, ic in (("k2", 1), ("kN", Nk - 1)):
        vk = e.geom.vk(ic)
        np.testing.assert_allclose(e.tb_hamiltonian(vk), f["zH_TB_" + tag], rtol=0, atol=1e-12)
        np.testing.assert_allclose(e.hamiltonian_hf(vk, f["Potential0"], f["rho0"]),
                                   f["zH_HF_" + tag], rtol=0, atol=1e-12)


@pytest.mark.parametrize("n", [1, 2, 3, 31, 32, 33, 64, 65, 100, 257, 600])
def test_zdiagred_random_hermitian(n):
    say("zdiagred random", n)
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = A + A.conj().T
    iu = max(1, n // 2 + 3) if n > 3 else n
    iu = min(iu, n)
    w, Z, info = zdiagred(np.tril(A), 1, iu)
    assert info == 0
    wr = np.linalg.eigvalsh(A)[:iu]
    np.testing.assert_allclose(w, wr, rtol=0, atol=1e-12 * max(1.0, abs(wr).max()) * n)
    assert np.abs(A @ Z - Z * w).max() < 1e-11 * n
    assert np.abs(Z.conj().T @ Z - np.eye(iu)).max() < 1e-12 * n


def test_zdiagred_large_merges():
    """n = 4200: the divide-and-conquer merges of order >= 2048 (the top two
    levels) multiply on the matrix-core GEMM path; eigenvalues against
    LAPACK, residual and orthogonality of the eigenvectors."""
    n = 4200
    say("zdiagred large merges", n)
    rng = np.random.default_rng(4200)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = A + A.conj().T
    iu = n // 2
    w, Z, info = zdiagred(np.tril(A), 1, iu)
    assert info == 0
    wr = np.linalg.eigvalsh(A)[:iu]
    scale = abs(wr).max()
    np.testing.assert_allclose(w, wr, rtol=0, atol=1e-13 * scale * np.sqrt(n))
    assert np.abs(A @ Z - Z * w).max() < 1e-14 * scale * n
    assert np.abs(Z.conj().T @ Z - np.eye(iu)).max() < 1e-13 * n


@pytest.mark.parametrize("case", ["t1k3", "t2k3", "t3k6"])
def test_zdiagred_graphene_matches_reference(case):
    say("zdiagred graphene", case)
    f = load_case(case)
    e = engine_for(case)
    H = f["zH_HF_k2"]
    iu = e.p.nUpper
    w, Z, info = zdiagred(H, 1, iu)
    assert info == 0
    np.testing.assert_allclose(w, f["eig_HF_k2"], rtol=0, atol=1e-10)
    Hf = full(H)
    assert np.abs(Hf @ Z - Z * w).max() < 1e-10
    # Gamma of the tight-binding model: exact symmetry degeneracies
    H0 = e.tb_hamiltonian(e.geom.vk(0))
    w0, Z0, info = zdiagred(H0, 1, e.n)
    np.testing.assert_allclose(w0, np.linalg.eigvalsh(full(H0)), rtol=0, atol=1e-10)
    assert np.abs(Z0.conj().T @ Z0 - np.eye(e.n)).max() < 1e-11


@pytest.mark.parametrize("case", ["t1k3", "t2k3", "t3k6"])
def test_solve_and_getfock_match_reference(case):
    say("solve+getfock", case)
    f = load_case(case)
    e = engine_for(case)
    bulk, bands, vecs = e.solve(f["rho0"], f["Potential0"])
    np.testing.assert_allclose(bands, f["Bands1"], rtol=0, atol=1e-10)
    # gauge-invariant: bulk + window density of the same states
    from oracle.pyoracle import System
    c = case_params(f)
    s = System(c["ntheta"], c["numk"], c["nfilling"], nlayers=c["nlayers"], xi_nm=c["xi_nm"])
    S, zS, nSM, _ = s.sort(bands, vecs)
    nF, EF, nOcc, nPart, Deg = s.fermi(S)
    ref = f["fermi1"]
    assert (nF, nOcc, nPart) == (int(ref[0]), int(ref[2]), int(ref[3]))
    win = e.getfock(np.zeros_like(bulk), zS, nSM, Deg, nOcc, nPart)
    rho_new = (win + bulk) / e.Nk
    np.testing.assert_allclose(rho_new, f["rhoNew1"], rtol=0, atol=1e-9)


def test_getfock_reproduces_degenerate_shell_quirk():
    say("getfock quirk t1k3")
    f = load_case("t1k3")
    e = engine_for("t1k3")
    from oracle.pyoracle import System
    s = System(1, 3, 1.0)
    zero = np.zeros((e.n, e.n, e.nnc), complex, order="F")
    bulk, bands, vecs, _ = s.solve(zero, np.zeros(e.n))
    S, zS, nSM, _ = s.sort(bands, vecs)
    nF, EF, nOcc, nPart, Deg = s.fermi(S)
    assert nPart == 2
    win = e.getfock(np.zeros_like(bulk), zS, nSM, Deg, nOcc, nPart)
    np.testing.assert_allclose(win, f["zFockWin0"], rtol=0, atol=1e-9)


@pytest.mark.parametrize("case", ["t1k3", "t2k3", "t3k6"])
def test_energies_match_reference(case):
    say("energies", case)
    f = load_case(case)
    e = engine_for(case)
    rho0 = f["rho0"]
    dens = np.real(np.diag(rho0[:, :, 0])) - 0.5
    out = e.energies(rho0, dens)
    np.testing.assert_allclose(out, f["energies0"][:4], rtol=1e-12, atol=1e-10)
    assert e.total_energy(out) == pytest.approx(f["energies0"][4], abs=1e-9)
    # GetHartreeHubbardEnergy with the spin-singlet call of Main.f90:307 (dens1 = dens2)
    hh = e.hartree_hubbard(dens, dens)
    np.testing.assert_allclose(hh, f["energies0"][2:4], rtol=1e-12, atol=1e-10)
    # unequal densities: 1/2 (d1+d2).L.(d1+d2) and d1.d2 against the reference's LongRange
    rng = np.random.default_rng(7)
    d2 = dens + 1e-3 * rng.standard_normal(dens.shape)
    hh = e.hartree_hubbard(dens, d2)
    Lr = f["LongRange"]
    s = dens + d2
    np.testing.assert_allclose(hh, [0.5 * s @ Lr @ s, dens @ d2], rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("case", ["t1k3", "t1k3s", "t2k3", "t2k3s", "t2k3xi10", "t3k6", "tri1k3",
                                  "t5k6s"])
def test_scf_trajectory_matches_reference(case):
    say("scf trajectory", case)
    f = load_case(case)
    c = case_params(f)
    e = engine_for(case) if not c["seeded"] else Engine(Setup(
        ntheta=c["ntheta"], nlayers=c["nlayers"], numk=c["numk"], nfilling=c["nfilling"],
        xi_nm=c["xi_nm"]))
    e.set_options()
    start = seeded_rho(e.n, e.nnc, c["nfilling"]) if c["seeded"] else None
    e.scf_init(start)
    traj = e.scf_run(c["itmax"])
    E, Cv = load_traj(case)
    assert len(traj) == len(E)
    np.testing.assert_allclose(traj[:, 1:6], E[:, 1:6], rtol=0, atol=1e-8)
    np.testing.assert_allclose(traj[:, 6], Cv[:, 1], rtol=0, atol=2e-6)
    np.testing.assert_allclose(traj[:, 7], Cv[:, 2], rtol=0, atol=1e-3)
    np.testing.assert_allclose(traj[:, 8], Cv[:, 3], rtol=0, atol=1e-8)


@pytest.mark.parametrize("case,occ,mix", [("t2k3", 2, 1), ("t2k3", 1, 0), ("t3k6", 2, 1)])
def test_scf_extension_modes_match_oracle(case, occ, mix):
    """Fermi-Dirac occupations / linear mixing have no reference counterpart:
    the GPU loop is checked against the oracle's restatement of the same modes."""
    say("scf modes", case, occ, mix)
    from oracle.pyoracle import System
    f = load_case(case)
    c = case_params(f)
    e = engine_for(case)
    e.set_options(occ_mode=occ, mix_mode=mix, kT=0.001, beta=0.3)
    start = seeded_rho(e.n, e.nnc, c["nfilling"])
    e.scf_init(start)
    traj = e.scf_run(6)
    s = System(c["ntheta"], c["numk"], c["nfilling"], nlayers=c["nlayers"], xi_nm=c["xi_nm"])
    its, ref, rho_ref, _ = s.scf(6, rho_start=start, occ_mode=occ, kT=0.001, mix_mode=mix, beta=0.3)
    np.testing.assert_allclose(traj[:, 1:9], ref[:, 1:9], rtol=0, atol=1e-8)
    np.testing.assert_allclose(e.density(), rho_ref, rtol=0, atol=1e-9)
    e.set_options()


@pytest.mark.parametrize("case", ["t5k6s", "t2k3"])
def test_mixed_precision_matches_fp64(case):
    """precision 1 (FP32 copy streamed by the Householder hemv while |dE| >
    switch_de, FP64 afterwards) against the FP64 path.
    (a) one iteration from the seeded start: energies within 1e-6 eV of FP64,
        Fermi level within 1e-6 eV;
    (b) four mixed-phase iterations started at the FP64 fixed point stay on it:
        total energy within 1e-6 eV, density within 1e-3 (Frobenius over all
        n^2 x 4 elements, an RMS below 1.5e-6 per element at n = 364; the
        broken-symmetry state has soft directions along which the rounding
        moves the density at no energy cost, measured 5e-5 to 2e-4);
    (c) a full solve from the seeded start switches to FP64 and converges
        (|dE| < 1e-8 on an FP64 iteration).  Its final state is reported, not
        compared: at integer filling several nearly degenerate symmetry-broken
        states are fixed points, and which one a seeded start reaches depends on
        perturbations of the size of the rounding."""
    say("mixed precision", case)
    f = load_case(case)
    c = case_params(f)
    e = engine_for(case)
    start = seeded_rho(e.n, e.nnc, c["nfilling"])
    opts = dict(occ_mode=2, mix_mode=1, kT=0.001, beta=0.3)
    e.set_options(**opts, etol=1e-8)
    e.scf_init(start)
    t64 = e.scf_run(300)
    rho64 = e.density()
    E64 = t64[-1, 1] - t64[-1, 8]
    assert abs(t64[-1, 8]) < 1e-8
    # (a)
    e.set_options(**opts, precision=1, switch_de=1e-14)
    e.scf_init(start)
    t1 = e.scf_run(1)
    d1 = np.abs(t1[0, 1:9] - t64[0, 1:9]).max()
    dEF = abs(t1[0, 9] - t64[0, 9])
    # (b)
    e.scf_init(rho64)
    tfx = e.scf_run(4)
    nfx = int(e.timings()[12])
    dfx = np.abs((tfx[:, 1] - tfx[:, 8]) - E64).max()
    drho = np.linalg.norm(e.density() - rho64)
    # (c)
    e.set_options(**opts, etol=1e-8, precision=1, switch_de=1e-4)
    e.scf_init(start)
    tmx = e.scf_run(300)
    nmixed = int(e.timings()[12])
    e.set_options()
    Emx = tmx[-1, 1] - tmx[-1, 8]
    say(f"  (a) first iteration max|d| {d1:.2e} eV, dE_F {dEF:.2e} eV; (b) at the fixed point "
        f"max|E - E64| {dfx:.2e} eV, |drho|_F {drho:.2e}; (c) iterations fp64 {len(t64)}, "
        f"mixed {len(tmx)} ({nmixed} in the mixed phase), final E - E64 {Emx - E64:+.3e} eV")
    assert d1 < 1e-6 and dEF < 1e-6
    assert nfx == 4 and dfx < 1e-6 and drho < 1e-3
    assert 0 < nmixed < len(tmx) and abs(tmx[-1, 8]) < 1e-8


def test_hamiltonian_calls_leave_the_scf_state_intact():
    """forge_tb_hamiltonian / forge_hamiltonian_hf between two forge_scf_run
    calls: the loop continues from its own state (site density, potential,
    energies), so scf_run(2) + Hamiltonians + scf_run(2) equals scf_run(4)."""
    say("hamiltonian between scf_run calls")
    e = engine_for("t2k3")
    e.set_options(occ_mode=2, mix_mode=1, kT=0.001, beta=0.3)
    start = seeded_rho(e.n, e.nnc, 1.0)
    e.scf_init(start)
    ref = e.scf_run(4)
    rho_ref = e.density()
    e.scf_init(start)
    a = e.scf_run(2)
    pot = np.linspace(-0.3, 0.2, e.n)
    e.tb_hamiltonian(np.array([0.1, -0.2]))
    e.hamiltonian_hf(np.array([0.05, 0.3]), pot, start)
    b = e.scf_run(2)
    e.set_options()
    got = np.vstack([a, b])
    np.testing.assert_array_equal(got[:, 1:9], ref[:, 1:9])
    np.testing.assert_array_equal(e.density(), rho_ref)
