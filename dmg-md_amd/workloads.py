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


AVOGADRO = 6.02214076e23


def cube_for_density(mass_amu: float, g_per_cm3: float) -> float:
    """Edge (A) of the cube holding `mass_amu` of matter at the given density."""
    return (mass_amu / (AVOGADRO * g_per_cm3) * 1e24) ** (1.0 / 3.0)


def random_separated(n: int, L: float, dmin: float, rng, batch: int = 0) -> np.ndarray:
    """n points uniform in the periodic cube [0, L)^3, every pair at least
    dmin apart (minimum image): random sequential addition in batches, each
    candidate tested against the accepted points and the earlier survivors
    of its own batch."""
    from scipy.spatial import cKDTree
    batch = batch or max(1024, n // 8)
    pts = np.zeros((0, 3))
    while len(pts) < n:
        cand = rng.uniform(0.0, L, (batch, 3))
        if len(pts):
            d, _ = cKDTree(pts, boxsize=L).query(cand, k=1, distance_upper_bound=dmin)
            cand = cand[~(d < dmin)]
        if len(cand) > 1:
            pairs = cKDTree(cand, boxsize=L).query_pairs(dmin, output_type="ndarray")
            drop = np.zeros(len(cand), bool)
            # greedy in index order: a candidate is dropped when an earlier kept one is too close
            for a, b in pairs[np.lexsort((pairs[:, 0], pairs[:, 1]))]:
                if not drop[a]:
                    drop[b] = True
            cand = cand[~drop]
        pts = np.concatenate([pts, cand[:n - len(pts)]])
    return pts


def _random_rotations(n: int, rng) -> np.ndarray:
    """n rotation matrices uniform on SO(3) (normalised Gaussian quaternions)."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)


def liquid_water(n_molecules: int = 64000, seed: int = 1, dmin: float = 2.5):
    """Config: liquid water, n_molecules H2O at 1.0 g/cm^3 in a periodic cube
    (64,000 molecules: 192,000 atoms, L = 124.2 A).  O-H 0.96 A, H-O-H
    104.5 degrees, uniformly random orientations; the oxygens are random
    centres at least dmin apart.  Atoms per molecule O, H, H; types O = 0,
    H = 1; Maxwell-Boltzmann velocities at 300 K."""
    rng = np.random.default_rng(seed)
    L = cube_for_density(n_molecules * (MASS["O"] + 2 * MASS["H"]), 1.0)
    o = random_separated(n_molecules, L, dmin, rng)
    half = np.radians(104.5) / 2.0
    local = 0.96 * np.array([[np.sin(half), np.cos(half), 0.0], [-np.sin(half), np.cos(half), 0.0]])
    R = _random_rotations(n_molecules, rng)
    h = o[:, None, :] + np.einsum("mij,kj->mki", R, local)
    pos = np.concatenate([o[:, None, :], h], axis=1).reshape(-1, 3)
    box = Box.orthogonal(L, L, L)
    pos = wrap(pos, box)
    types = np.tile(np.array([0, 1, 1], np.int32), n_molecules)
    m = np.where(types == 0, MASS["O"], MASS["H"])
    return dict(box=box, pos=soa(pos), types=types, mass=m, vel=maxwell_boltzmann(m, 300.0, seed + 1))


def amorphous_carbon(n: int, seed: int = 1, dmin: float = 1.2):
    """A periodic amorphous-carbon melt of n atoms at 2.0 g/cm^3, random
    positions at least dmin apart, 5000 K velocities: the bulk of the domain
    share below, as a small periodic sample."""
    rng = np.random.default_rng(seed)
    L = cube_for_density(n * MASS["C"], 2.0)
    p = random_separated(n, L, dmin, rng)
    m = np.full(n, MASS["C"])
    return dict(box=Box.orthogonal(L, L, L), pos=soa(p), types=np.zeros(n, np.int32), mass=m,
                vel=maxwell_boltzmann(m, 5000.0, seed + 1))


def amorphous_carbon_share(n: int = 2048000, seed: int = 1, dmin: float = 1.2, shell: float = 5.0,
                           margin: float = 30.0):
    """Config: one GPU's share of an amorphous carbon melt: n atoms at
    2.0 g/cm^3 in the cube [0, L)^3 (2,048,000 atoms: L = 273.3 A), random
    positions at least dmin apart, 5000 K velocities, followed by a static
    ghost shell: the periodic images of the atoms within `shell` A outside
    the cube.  The share is evaluated in an open box enclosing cube and
    shell with `margin` A of room beyond the shell (owned atoms pushed apart
    by close pairs may leave the cube during a run); `owned` = n and
    `ghost_of[k]` is the owned atom that ghost k images; `origin` is where
    the cube starts in the box."""
    rng = np.random.default_rng(seed)
    L = cube_for_density(n * MASS["C"], 2.0)
    p = random_separated(n, L, dmin, rng)
    ghosts, owner = [], []
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for c in (-1, 0, 1):
                if a == b == c == 0:
                    continue
                q = p + L * np.array([a, b, c])
                keep = np.all((q > -shell) & (q < L + shell), axis=1)
                ghosts.append(q[keep])
                owner.append(np.nonzero(keep)[0])
    g = np.concatenate(ghosts)
    origin = shell + margin
    pos = np.concatenate([p, g]) + origin  # open box [0, L + 2 origin)^3
    E = L + 2 * origin
    m = np.full(len(pos), MASS["C"])
    vel = maxwell_boltzmann(m[:n], 5000.0, seed + 1)
    return dict(box=Box.orthogonal(E, E, E, pbc=(0, 0, 0)), pos=soa(pos), types=np.zeros(len(pos), np.int32),
                mass=m, vel=vel, owned=n, ghost_of=np.concatenate(owner), L=L, shell=shell, origin=origin)


# the serving stream's materials: lattice, constant (A), atoms per cubic
# cell, mass, and the NVE step (fs) of the material's MD configuration
SERVING_MATERIALS = {"carbon": (diamond, 3.567, 8, MASS["C"], 1.0),
                     "cu": (fcc, 3.615, 4, MASS["Cu"], 2.0)}


def vacancy_lattice(material: str, n: int, rng, vacancies: bool = True):
    """n^3 cubic cells of the material's lattice (SERVING_MATERIALS), 30% of
    the sites left vacant when `vacancies`, Gaussian jitter 0.05 A, 300 K
    Maxwell-Boltzmann velocities; dict(box, pos (3,N), types, mass, vel)."""
    lattice, a, per_cell, mass, dt = SERVING_MATERIALS[material]
    pos, box = lattice(n, a)
    if vacancies:
        keep = rng.random(pos.shape[0]) >= 0.3
        if keep.sum() < 2:
            keep[:2] = True
        pos = pos[keep]
    pos = wrap(pos + rng.normal(0.0, 0.05, pos.shape), box)
    N = pos.shape[0]
    m = np.full(N, mass)
    return dict(box=box, pos=soa(pos), types=np.zeros(N, np.int32), mass=m,
                vel=maxwell_boltzmann(m, 300.0, int(rng.integers(1 << 30))))


def serving_request(rng, material: str, n_min: int = 64, n_max: int = 20000, vacancy_share: float = 0.5):
    """One request of the serving stream: a cubic replica of the material's
    lattice whose site count is drawn log-uniform in [n_min, n_max] (the
    nearest cube of cells, at least n_min sites), Gaussian jitter 0.05 A,
    30% of the sites left vacant with probability vacancy_share, 300 K
    Maxwell-Boltzmann velocities.  Returns dict(box, pos (3,N), types, mass,
    vel, dt_fs)."""
    lattice, a, per_cell, mass, dt = SERVING_MATERIALS[material]
    target = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
    n = max(1, int(round((target / per_cell) ** (1.0 / 3.0))))
    while per_cell * n ** 3 < n_min:
        n += 1
    r = vacancy_lattice(material, n, rng, vacancies=rng.random() < vacancy_share)
    r.update(dt_fs=dt, material=material)
    return r


def serving_stream(count: int = 4096, seed: int = 7, n_min: int = 64, n_max: int = 20000, md_steps: int = 50):
    """The serving configuration: `count` requests, each a carbon or Cu
    request (serving_request) with equal odds; every second one a single
    point (steps 0), the others md_steps-step NVE segments at the material's
    step.  Returns a list of (request dict, model index 0 carbon / 1 Cu,
    steps)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        mat = "carbon" if rng.random() < 0.5 else "cu"
        r = serving_request(rng, mat, n_min, n_max)
        out.append((r, 0 if mat == "carbon" else 1, md_steps if k % 2 else 0))
    return out


def diamond_npt_model_xyz(cells=(20, 6, 6), seed: int = 301, scale: float = 1.0023) -> str:
    """model.xyz text of the reference's published NPT timing job (5,760-atom
    diamond carbon for cells 20x6x6): conventional cells of edge
    3.57 * scale A offset by 1 A, the diamond basis in the reference
    generator's order, and velocities drawn per atom as three Gaussians of
    sigma sqrt(k_B 300 K / m) / 10.18051 from Python's random.Random(seed)
    (tests/npt/generate_inputs.py:28-48 describes the job's input)."""
    import math
    import random
    basis = ((0, 0, 0), (.25, .25, .25), (.5, .5, 0), (.75, .75, .25),
             (.5, 0, .5), (.75, .25, .75), (0, .5, .5), (.25, .75, .75))
    nx, ny, nz = cells
    a = 3.57 * scale
    sigma = math.sqrt(K_B * 300.0 / MASS["C"]) / 10.18051
    rng = random.Random(seed)
    out = [str(8 * nx * ny * nz),
           'pbc="T T T" Lattice="%r 0 0 0 %r 0 0 0 %r" Properties=species:S:1:pos:R:3:mass:R:1:vel:R:3'
           % (3.57 * nx * scale, 3.57 * ny * scale, 3.57 * nz * scale)]
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                for x, y, z in basis:
                    p = (1 + a * (i + x), 1 + a * (j + y), 1 + a * (k + z))
                    v = [rng.gauss(0.0, sigma) for _ in range(3)]
                    out.append("C " + " ".join(format(c, ".17g") for c in (*p, MASS["C"], *v)))
    return "\n".join(out) + "\n"
