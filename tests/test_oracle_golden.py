"""The FP64 oracle against the reference's own pins: the GPUMD 5.9.1 golden
per-atom outputs (tests/baseline, tolerances of its manifest.json) and the
reference's compiled FP64 CPU evaluator (oracle/_ref, nep_cpu.cpp)."""
import json
import os

import numpy as np
import pytest

from conftest import BASELINE
from oracle import OracleNEP
from xyz_io import XYZ_VIRIAL_TO_SOA, read_xyz, types_of

CASES = {
    "single_small_static": ("main", "nep_C.txt"),
    "single_large_nve": ("main", "nep_C.txt"),
    "multi_nvt_restart": ("initial", "nep_water.txt"),
    "nep_zbl_boundary": ("main", "nep_BaTiO3_zbl.txt"),
}
TOL = json.load(open(os.path.join(BASELINE, "manifest.json")))["tolerances"]


def load_case(case):
    stage, pot = CASES[case]
    g = read_xyz(os.path.join(BASELINE, "goldens", case, stage, "initial.xyz"))
    path = os.path.join(BASELINE, "inputs", "potentials", pot)
    return g, path, types_of(g["species"], path)


def within(actual, expected, tol):
    return np.all(np.abs(actual - expected) <= tol["atol"] + tol["rtol"] * np.abs(expected))


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_matches_gpumd_goldens(case):
    g, path, t = load_case(case)
    o = OracleNEP(path).compute(g["h"], g["pbc"], t, g["pos"].T.copy())
    assert within(o["force"].T, g["forces"], TOL["force"])
    assert within(o["pe"], g["energy_atom"], TOL["energy"])
    assert within(o["virial"].T, g["virial"][:, XYZ_VIRIAL_TO_SOA], TOL["virial"])
    assert abs(o["pe"].sum() - g["energy_total"]) <= TOL["energy"]["atol"] + TOL["energy"]["rtol"] * abs(g["energy_total"]) * 8


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_matches_reference_cpu_evaluator(case):
    ref = pytest.importorskip("ref")
    if not ref.available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    g, path, t = load_case(case)
    pos = g["pos"].T.copy()
    o = OracleNEP(path).compute(g["h"], g["pbc"], t, pos)
    r = ref.RefNEP(path).compute(g["h"], g["pbc"], t, pos)
    for k in ("pe", "force", "virial"):
        scale = max(1.0, np.abs(r[k]).max())
        assert np.abs(o[k] - r[k]).max() <= 1e-12 * scale, k


def test_oracle_matches_reference_on_jittered_multitype():
    """A random 2-type water-like cluster in a periodic box (not a golden)."""
    ref = pytest.importorskip("ref")
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    path = os.path.join(BASELINE, "inputs", "potentials", "nep_water.txt")
    rng = np.random.default_rng(5)
    L = 13.0
    n = 60
    pos = rng.uniform(0, L, (3, n))
    # keep a 1 A minimum separation
    keep = [0]
    for i in range(1, n):
        d = pos[:, keep] - pos[:, i:i + 1]
        d -= L * np.round(d / L)
        if np.sqrt((d * d).sum(0)).min() > 1.0:
            keep.append(i)
    pos = np.ascontiguousarray(pos[:, keep])
    t = (np.arange(pos.shape[1]) % 3 != 0).astype(np.int32)
    h = np.diag([L, L, L])
    o = OracleNEP(path).compute(h, [1, 1, 1], t, pos)
    r = ref.RefNEP(path).compute(h, [1, 1, 1], t, pos)
    for k in ("pe", "force", "virial"):
        assert np.abs(o[k] - r[k]).max() <= 1e-11 * max(1.0, np.abs(r[k]).max()), k


def test_oracle_replays_gpumd_nvt_ber_trajectory():
    """The FP64 oracle's velocity-Verlet step with a Berendsen rescale
    (factor sqrt(1 + (T_target/T - 1)/tau), T over 3N degrees of freedom,
    target ramped linearly as t0 + (t1 - t0) * step/steps; reference
    ensemble_control.cpp Berendsen branch, integrate.cu) replays the two
    nvt_ber segments of multi_nvt_restart.  GPUMD's FP32 trajectory sits up to
    1.5e-8 A/fs from the replay, so velocities are held to 3e-8 A/fs (5.6x
    the manifest's bitwise-rerun floor); a target ramp shifted by one step
    moves velocities by about 1e-6 A/fs and fails."""
    from oracle import thermo, velocity_verlet
    from xyz_io import TIME_UNIT_CONVERSION as TU
    path = os.path.join(BASELINE, "inputs", "potentials", "nep_water.txt")
    model = read_xyz(os.path.join(BASELINE, "inputs", "multi_nvt_restart", "model.xyz"))
    gold_dir = os.path.join(BASELINE, "goldens", "multi_nvt_restart", "initial")
    mass = read_xyz(os.path.join(gold_dir, "restart.xyz"))["mass"].copy()
    t = types_of(model["species"], path)
    orc = OracleNEP(path, round_params_to_float=True)
    h = model["h"]
    pos = np.ascontiguousarray(model["pos"].T)
    vel = np.ascontiguousarray(model["vel"].T) * TU
    force = np.ascontiguousarray(orc.compute(h, model["pbc"], t, pos)["force"])
    frames = []
    dt = 0.1 / TU
    for t0, t1, steps, every in ((100.0, 120.0, 4, 2), (120.0, 130.0, 3, 1)):
        for step in range(steps):
            velocity_verlet(1, dt, mass, pos, vel, force)
            r = orc.compute(h, model["pbc"], t, pos)
            force = np.ascontiguousarray(r["force"])
            velocity_verlet(0, dt, mass, pos, vel, force)
            temp = thermo(h, mass, r["pe"], vel, r["virial"])[0]
            target = t0 + (t1 - t0) * step / steps
            vel *= np.sqrt(1.0 + (target / temp - 1.0) / 100.0)
            if (step + 1) % every == 0:
                frames.append((pos.T.copy(), vel.T / TU))
    vel_tol = dict(TOL["velocity"], atol=3e-8)
    for k, (p, v) in enumerate(frames):
        g = read_xyz(os.path.join(gold_dir, "trajectory.xyz"), frame=k)
        assert within(v, g["vel"], vel_tol), (k, np.abs(v - g["vel"]).max())
        assert within(p, g["pos"], TOL["position"]), (k, np.abs(p - g["pos"]).max())


COVERAGE_MODELS = ["nep5_ch.txt", "nep4_zbl_flexible_ch.txt", "nep5_zbl_typewise_ch.txt", "nep4_big_ch.txt"]


def coverage_sample(seed=7, n=100, L=11.0, dmin=0.5):
    """Random 2-type (C, H) periodic configuration; dmin 0.5 A puts pairs
    inside every ZBL range of the coverage models."""
    rng = np.random.default_rng(seed)
    pos = []
    while len(pos) < n:
        x = rng.uniform(0, L, 3)
        if all(np.linalg.norm((x - y) - L * np.round((x - y) / L)) > dmin for y in pos):
            pos.append(x)
    pos = np.ascontiguousarray(np.array(pos).T)
    return np.diag([L, L, L]), pos, (np.arange(n) % 3 == 0).astype(np.int32)


@pytest.mark.parametrize("name", COVERAGE_MODELS)
def test_oracle_matches_reference_on_nep5_and_zbl_forms(name):
    """NEP5 (typewise output bias), the flexible per-pair ZBL table and the
    typewise ZBL cutoff: no reference golden covers them, so the oracle is
    pinned to the reference's own FP64 evaluator (nep_cpu.cpp) on a random
    2-type system with short pairs."""
    from conftest import POTENTIALS
    ref = pytest.importorskip("ref")
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    path = os.path.join(POTENTIALS, name)
    h, pos, t = coverage_sample()
    d = pos[:, :, None] - pos[:, None, :]
    d -= h[0, 0] * np.round(d / h[0, 0])
    r = np.sqrt((d * d).sum(0))[np.triu_indices(pos.shape[1], 1)]
    assert (r < 0.9).sum() > 0  # pairs inside the typewise ZBL cutoffs
    o = OracleNEP(path).compute(h, [1, 1, 1], t, pos)
    rr = ref.RefNEP(path).compute(h, [1, 1, 1], t, pos)
    for k in ("pe", "force", "virial"):
        scale = max(1.0, np.abs(rr[k]).max())
        assert np.abs(o[k] - rr[k]).max() <= 1e-12 * scale, (k, np.abs(o[k] - rr[k]).max(), scale)
