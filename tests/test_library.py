"""CPU checks of the product: the C-ABI library loads and exports every
symbol include/forge_hf.h declares, and the host-side geometry and Setup
parameters reproduce the reference (no device calls)."""
import ctypes
import os

import numpy as np
import pytest

from golden_io import case_params, load_case
from forge_amd import lib as L
from forge_amd.geometry import build_geometry, ndim_of, number_neighbor_cells, order_neighbor_cells
from forge_amd.hf import Setup


@pytest.fixture(scope="module")
def built_lib():
    if not os.path.exists(L.LIB_PATH):
        from forge_amd.build import build
        build()
    return ctypes.CDLL(L.LIB_PATH)


def test_library_exports_every_declared_symbol(built_lib):
    names = L.header_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(built_lib, n)]
    assert not missing, missing
    assert set(names) == set(L._SIGS)


def test_library_is_gfx950_and_reports_devices_without_computing(built_lib):
    lib = L.load()
    assert lib.forge_device_count() >= 0
    # error text is reachable without a device
    assert isinstance(L.last_error(), str)


def test_ctx_create_rejects_bad_params_before_any_device_work():
    lib = L.load()
    p = Setup(ntheta=1, numk=3).params(4)
    p.fphase = 1
    ctx = ctypes.c_void_p()
    z = np.zeros(64)
    i = np.zeros(64, np.int64)
    rc = lib.forge_ctx_create(ctypes.byref(p), z.ctypes.data, z.ctypes.data, z.ctypes.data,
                              i.ctypes.data, i.ctypes.data, i.ctypes.data, ctypes.byref(ctx))
    assert rc < 0 and "fphase" in L.last_error()


@pytest.mark.parametrize("nfilling", [40.0, -40.0])
def test_ctx_create_rejects_a_filling_outside_the_band_window(nfilling):
    # nF = nint(Nk (ndim/2 - nLower + 1) + Nk/2 nu) must leave a state above
    # and below the Fermi level inside the window (Main.f90:400-401)
    lib = L.load()
    p = Setup(ntheta=1, numk=3, nfilling=nfilling).params(4)
    ctx = ctypes.c_void_p()
    z = np.zeros(64)
    i = np.zeros(64, np.int64)
    rc = lib.forge_ctx_create(ctypes.byref(p), z.ctypes.data, z.ctypes.data, z.ctypes.data,
                              i.ctypes.data, i.ctypes.data, i.ctypes.data, ctypes.byref(ctx))
    assert rc == -2 and "band window" in L.last_error()


@pytest.mark.parametrize("case", ["t1k3", "t2k3", "t3k6", "tri1k3"])
def test_product_geometry_is_bitwise_reference(case):
    f = load_case(case)
    c = case_params(f)
    g = build_geometry(c["ntheta"], c["nlayers"], c["numk"])
    assert np.array_equal(g.Coords, f["Coords"])
    assert np.array_equal(g.tn, f["tn"])
    assert np.array_equal(g.MomentaValues, f["MomentaValues"])
    assert np.array_equal(g.nMomentaComponents, f["nMomentaComponents"])
    assert np.array_equal(g.nUnitCell_1, f["nUnitCell_1"])
    assert np.array_equal(g.nUnitCell_2, f["nUnitCell_2"])


def test_large_cells_have_the_configured_sizes():
    assert ndim_of(5, 2) == 364 and ndim_of(20, 2) == 5044 and ndim_of(31, 2) == 11908
    assert ndim_of(31, 3) == 17862
    assert build_geometry(31, 2, 3).ndim == 11908
    assert [ndim_of(m, 2) for m in range(8, 15)] == [868, 1084, 1324, 1588, 1876, 2188, 2524]


def test_neighbor_cells_match_reference_counts():
    assert number_neighbor_cells(1) == 4
    u1, u2 = order_neighbor_cells(1, 4)
    assert list(u1) == [0, -1, -1, 0] and list(u2) == [0, 0, 1, -1]
    assert number_neighbor_cells(2) == 10


def test_setup_band_window_matches_reference_params():
    f = load_case("t1k3")
    prm = f["params"]  # ndim, numk, nfilling, xi, nLower, nUpper, nnc, nlayers
    p = Setup(ntheta=1, numk=3, nfilling=1.0).params(4)
    assert (p.ndim, p.nLower, p.nUpper) == (int(prm[0]), int(prm[4]), int(prm[5]))
    assert p.xi == pytest.approx(prm[3], rel=1e-15)


@pytest.mark.parametrize("ntheta,nlayers", [(1, 2), (2, 2), (1, 3), (3, 2)])
def test_c3_pairs_host_matches_oracle(ntheta, nlayers):
    """forge_c3_pairs (C3_RelatedPoints, host code of the library) against the
    oracle's restatement, which the reference's C3 trajectories pin."""
    from oracle.pyoracle import System
    s = System(ntheta, 3, 1.0, nlayers=nlayers)
    perm, cnt = s.c3_pairs()
    g = build_geometry(ntheta, nlayers, 3)
    out = np.zeros(g.ndim, np.int64)
    got = L.load().forge_c3_pairs(g.Coords.ctypes.data, g.ndim, nlayers, out.ctypes.data)
    assert got == cnt == g.ndim
    assert np.array_equal(out - 1, perm)
    assert list(g.RotateLayers) == list(s.rot)


def test_c3_pairs_rejects_an_odd_layer_split():
    out = np.zeros(27, np.int64)
    z = np.zeros((27, 3), order="F")
    assert L.load().forge_c3_pairs(z.ctypes.data, 27, 2, out.ctypes.data) < 0


@pytest.mark.parametrize("ntheta,nlayers", [(1, 2), (2, 2), (1, 3)])
def test_c2_pairs_host_matches_oracle(ntheta, nlayers):
    """forge_c2_pairs (C2_RelatedPoints, host code of the library) against the
    oracle's, which the reference's symmetrised trajectories pin."""
    from oracle.pyoracle import System
    s = System(ntheta, 3, 1.0, nlayers=nlayers)
    perm, cnt = s.c2_pairs()
    g = build_geometry(ntheta, nlayers, 3)
    out = np.zeros(g.ndim, np.int64)
    got = L.load().forge_c2_pairs(g.Coords.ctypes.data, g.ndim, nlayers, out.ctypes.data)
    assert got == cnt == g.ndim // 2
    assert np.array_equal(out - 1, perm)


@pytest.mark.parametrize("ntheta", [2, 3, 5])
def test_relaxed_geometry_is_bitwise_reference(ntheta):
    """nrelax = 1: the product's LatticeRelaxation (Geometry.f90:279-1135)
    against the cell the reference's own routine returns
    (oracle/ref_relax.f90), bit for bit."""
    f = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             f"relax_t{ntheta}.npz"))
    assert np.array_equal(build_geometry(ntheta, 2, 3).Coords, f["Coords"])
    g = build_geometry(ntheta, 2, 3, nrelax=1)
    assert np.array_equal(g.Coords, f["CoordsRelaxed"])
    assert np.abs(g.Coords - f["Coords"]).max() > 1e-4


def test_relaxation_table_holds_the_reference_binary32_values():
    """The table covers ntheta 2..35 with ten (n, m) entries each, and every
    component is a binary32 value (the reference's literals are
    default-real constants): storing the decimal text as a double would
    move the relaxed sites by about 1e-10 and fail the bitwise test above."""
    from forge_amd.geometry import _RELAX_TABLE
    rows = [l.split() for l in open(_RELAX_TABLE) if l.strip() and not l.startswith("#")]
    cells = sorted({int(r[0]) for r in rows})
    assert cells == list(range(2, 36))
    assert all(sum(1 for r in rows if int(r[0]) == t) == 10 for t in cells)
    for r in rows:
        for v in (float(r[3]), float(r[4])):
            assert float(np.float32(v)) == v
