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
