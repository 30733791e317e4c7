"""Seeded initial densities shared by every backend (host side).

The start state is rho_0(cell 0) = diag(1/2 + nu/(2 n) + sigma * (g - mean g)),
with every off-diagonal element and every other stored cell zero, where g is a
vector of standard normal draws.  The draws come from a portable generator
(SplitMix64 -> 53-bit uniforms -> Box-Muller) so that Python, C and Fortran
hosts produce bit-identical starts from the same seed.  The reference has no
seeded start; its initial state is either the non-interacting ground state
(Main.f90:236-288) or a density checkpoint (Main.f90:187-235), and the seeded
state reaches it through the checkpoint format (forge_amd.checkpoint).
"""
from __future__ import annotations

import math

import numpy as np

MASK64 = (1 << 64) - 1
DEFAULT_SEED = 20261203


def splitmix64_stream(seed: int, count: int) -> np.ndarray:
    """`count` consecutive SplitMix64 outputs (uint64) starting from `seed`."""
    out = np.empty(count, dtype=np.uint64)
    s = seed & MASK64
    for i in range(count):
        s = (s + 0x9E3779B97F4A7C15) & MASK64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        out[i] = z ^ (z >> 31)
    return out


def standard_normals(seed: int, count: int) -> np.ndarray:
    """Box-Muller normals: pairs (u1, u2) of 53-bit uniforms in (0, 1]."""
    npairs = (count + 1) // 2
    raw = splitmix64_stream(seed, 2 * npairs)
    u = ((raw >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    g = np.empty(2 * npairs)
    g[0::2] = r * np.cos(2.0 * math.pi * u2)
    g[1::2] = r * np.sin(2.0 * math.pi * u2)
    return g[:count]


def seeded_site_density(ndim: int, nfilling: float, sigma: float = 1e-3,
                        seed: int = DEFAULT_SEED) -> np.ndarray:
    """Per-site occupation (one spin) of the seeded uniform + Gaussian start."""
    g = standard_normals(seed, ndim)
    g = g - g.mean()
    return 0.5 + nfilling / (2.0 * ndim) + sigma * g


def seeded_rho(ndim: int, ncells: int, nfilling: float, sigma: float = 1e-3,
               seed: int = DEFAULT_SEED) -> np.ndarray:
    """Real-space density matrix rho(i, j, cell) in Fortran order (complex128)."""
    rho = np.zeros((ndim, ndim, ncells), dtype=np.complex128, order="F")
    d = seeded_site_density(ndim, nfilling, sigma, seed)
    rho[np.arange(ndim), np.arange(ndim), 0] = d
    return rho


def seeded_rho_2spins(ndim: int, ncells: int, nfilling: float, sigma: float = 0.02,
                      seed: int = DEFAULT_SEED) -> list:
    """Spin-split start for numS = 2: the seeded start of each spin with its
    own seed (seed, seed + 1) and a larger spread, so the two spins differ."""
    return [seeded_rho(ndim, ncells, nfilling, sigma=sigma, seed=seed + s) for s in (0, 1)]
