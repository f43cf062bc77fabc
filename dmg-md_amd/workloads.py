"""Synthetic, seeded atomistic systems for the benchmark and tests.

Lattices, Gaussian jitter and Maxwell-Boltzmann velocities in the units of the
reference engine: positions in A, velocities in A per natural time unit
(fs / 10.18051, src/model_parser.cpp:397-400), masses in amu.
"""
from __future__ import annotations

import numpy as np

from .nep import K_B, Box

# standard atomic weights (amu) used as the default masses of model.xyz
MASS = {"H": 1.008, "C": 12.011, "O": 15.999, "Ti": 47.867, "Cu": 63.546, "Ba": 137.327}


def fcc(n: int, a: float):
    """n x n x n conventional fcc cells; returns (positions[N,3], Box)."""
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    return _lattice(n, a, basis)


def diamond(n: int, a: float):
    """n x n x n conventional diamond cells (8 atoms per cell)."""
    f = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    basis = np.concatenate([f, f + 0.25])
    return _lattice(n, a, basis)


def _lattice(n, a, basis):
    g = np.arange(n)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a
    return pos, Box.orthogonal(n * a, n * a, n * a)


def jitter(pos: np.ndarray, sigma: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return pos + rng.normal(0.0, sigma, pos.shape)


def maxwell_boltzmann(masses: np.ndarray, temperature: float, seed: int) -> np.ndarray:
    """Velocities (3, N) at exactly `temperature` with zero total momentum."""
    rng = np.random.default_rng(seed)
    N = masses.shape[0]
    v = rng.normal(0.0, 1.0, (3, N)) * np.sqrt(K_B * temperature / masses)[None, :]
    p = (v * masses[None, :]).sum(1) / masses.sum()
    v -= p[:, None]
    t_now = (masses[None, :] * v * v).sum() / (3 * N * K_B)
    return v * np.sqrt(temperature / t_now)


def wrap(pos: np.ndarray, box: Box) -> np.ndarray:
    s = np.linalg.solve(box.h, pos.T)
    s -= np.floor(s)
    return (box.h @ s).T


def soa(pos: np.ndarray) -> np.ndarray:
    """(N,3) -> structure-of-arrays (3,N) as the device layout expects."""
    return np.ascontiguousarray(pos.T)


def carbon_diamond_4096(seed: int = 1):
    """Config: diamond 8x8x8 (4096 C), a=3.567 A, jitter 0.05 A, 300 K."""
    pos, box = diamond(8, 3.567)
    pos = wrap(jitter(pos, 0.05, seed), box)
    m = np.full(pos.shape[0], MASS["C"])
    return dict(box=box, pos=soa(pos), types=np.zeros(pos.shape[0], np.int32), mass=m,
                vel=maxwell_boltzmann(m, 300.0, seed + 1))


def copper_fcc(n: int = 40, seed: int = 1):
    """Config: Cu fcc n^3 cells (4 n^3 atoms), a=3.615 A, jitter 0.03 A, 300 K."""
    pos, box = fcc(n, 3.615)
    pos = wrap(jitter(pos, 0.03, seed), box)
    m = np.full(pos.shape[0], MASS["Cu"])
    return dict(box=box, pos=soa(pos), types=np.zeros(pos.shape[0], np.int32), mass=m,
                vel=maxwell_boltzmann(m, 300.0, seed + 1))
