"""ctypes binding of libforge_hip.so (include/forge_hf.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libforge_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "forge_hf.h")

i64 = C.c_int64
i32 = C.c_int32
dbl = C.c_double
vp = C.c_void_p


class ForgeError(RuntimeError):
    pass


class ForgeParams(C.Structure):
    _fields_ = [(n, i64) for n in ("ndim", "nlayers", "numk", "numI", "numC", "nscreen",
                                   "fphase", "nnc", "nLower", "nUpper", "numb")] + \
               [(n, dbl) for n in ("nfilling", "U", "alpha", "alphaH", "Delta", "xi", "tz",
                                   "a0", "r0", "d0", "energy_tol", "step_alt")]


class ScfOptions(C.Structure):
    _fields_ = [("occ_mode", i32), ("mix_mode", i32), ("kT", dbl), ("beta", dbl),
                ("etol", dbl), ("precision", i32), ("timing", i32), ("switch_de", dbl),
                ("graph", i32), ("reduction", i32)]


_SIGS = {
    "forge_last_error": (C.c_char_p, []),
    "forge_device_count": (C.c_int, []),
    "forge_ctx_create": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, C.POINTER(vp)]),
    "forge_ctx_destroy": (C.c_int, [vp]),
    "forge_longrange": (C.c_int, [vp, vp]),
    "forge_tb_hamiltonian": (C.c_int, [vp, vp, vp]),
    "forge_hamiltonian_hf": (C.c_int, [vp, vp, vp, vp, vp]),
    "forge_zdiagred": (C.c_int, [vp, i64, vp, i64, i64, C.POINTER(i32)]),
    "forge_zdiagred_noalloc": (C.c_int, [vp, i64, vp, i64, i64, C.POINTER(i32)]),
    "forge_z_vector_mul": (C.c_int, [vp, vp, i64, i64, vp]),
    "forge_z_matrix_mul": (C.c_int, [vp, vp, vp, i64, i64, i64]),
    "forge_z_matrix_overlap": (C.c_int, [vp, vp, vp, i64, i64, vp]),
    "forge_c_matrix_mul": (C.c_int, [vp, vp, vp, i64, i64, i64]),
    "forge_solve": (C.c_int, [vp, vp, vp, vp, vp, vp]),
    "forge_getfock": (C.c_int, [vp, vp, vp, vp, dbl, i64, i64]),
    "forge_energies": (C.c_int, [vp, vp, vp, vp]),
    "forge_hartree_hubbard": (C.c_int, [vp, vp, vp, vp]),
    "forge_scf_set_options": (C.c_int, [vp, vp]),
    "forge_c3_pairs": (i64, [vp, i64, i64, vp]),
    "forge_solve_c3": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp]),
    "forge_getfock_c3": (C.c_int, [vp, vp, vp, vp, dbl, i64, i64, vp, vp]),
    "forge_scf_set_c3": (C.c_int, [vp, vp, vp]),
    "forge_c2_pairs": (i64, [vp, i64, i64, vp]),
    "forge_symmetrize": (C.c_int, [vp, vp, i32, vp]),
    "forge_scf_set_symmetry": (C.c_int, [vp, i32, vp]),
    "forge_valley_phase": (C.c_int, [vp, vp, vp, vp]),
    "forge_valley_transform": (C.c_int, [vp, vp, vp]),
    "forge_scf_set_valley": (C.c_int, [vp, vp]),
    "forge_scf_init": (C.c_int, [vp, vp]),
    "forge_scf_init_guess": (C.c_int, [vp, vp]),
    "forge_scf_init_guess_update": (C.c_int, [vp, vp, i64, vp, vp, vp, vp]),
    "forge_scf_init_update": (C.c_int, [vp, vp, i64, vp, vp, vp, vp]),
    "forge_scf_run": (C.c_int, [vp, i64, vp, C.POINTER(i64)]),
    "forge_scf_get_density": (C.c_int, [vp, vp]),
    "forge_scf_timings": (C.c_int, [vp, vp]),
    "forge_scf2_create": (C.c_int, [vp, vp, C.POINTER(vp)]),
    "forge_scf2_destroy": (C.c_int, [vp]),
    "forge_scf2_init": (C.c_int, [vp, vp]),
    "forge_scf2_run": (C.c_int, [vp, i64, vp, C.POINTER(i64)]),
    "forge_scf2_get_density": (C.c_int, [vp, vp]),
    "forge_sweep_create": (C.c_int, [vp, i64, C.POINTER(vp)]),
    "forge_sweep_destroy": (C.c_int, [vp]),
    "forge_sweep_run": (C.c_int, [vp, i64, vp, vp]),
    "forge_mlp_param_count": (i64, [i64, i64]),
    "forge_mlp_loss_grad": (C.c_int, [vp, i64, i64, vp, vp, vp, i64, vp, vp]),
    "forge_mlp_train": (C.c_int, [vp, i64, i64, vp, vp, vp, i64, i64, i64, dbl, C.c_uint64, vp]),
    "forge_mlp_infer": (C.c_int, [vp, i64, i64, vp, i64, i32, vp]),
}

_lib = None


def header_symbols():
    """Function names declared in include/forge_hf.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(forge_\w+)\s*\(", txt, re.M)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ForgeError(f"{LIB_PATH} is not built; run `python -m forge_amd.build`")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error() -> str:
    return load().forge_last_error().decode()


def check(rc: int, what: str) -> int:
    if rc < 0:
        raise ForgeError(f"{what} failed ({rc}): {last_error()}")
    return rc


def device_count() -> int:
    return load().forge_device_count()
