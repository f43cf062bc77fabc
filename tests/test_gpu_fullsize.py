"""The benchmark configuration at its full size: 256,000-atom Cu fcc with the
c1 NEP4 model.  Forces, energies and virials of every atom against the FP64
oracle (FP32 and BF16-operand ANN), and an NVE run whose total energy per
atom stays within 1e-5 eV of its start (the reference's long-NVE metrics:
energy per atom, its largest excursion and its linear drift,
tests/long_nve/long_nve_common.py:1016-1047)."""
import os

import numpy as np
import pytest

from conftest import POTENTIALS
from gpu_helpers import errors, gpu_compute
from oracle import OracleNEP

pytestmark = pytest.mark.gpu

C1 = os.path.join(POTENTIALS, "nep_c1_cu.txt")
K_B = 8.617343e-5


@pytest.fixture(scope="module")
def cu256k(workloads):
    return workloads.copper_fcc(n=40, seed=1)


@pytest.fixture(scope="module")
def oracle_cu256k(cu256k):
    s = cu256k
    return OracleNEP(C1, round_params_to_float=True).compute(s["box"].h, s["box"].pbc, s["types"], s["pos"])


def test_fullsize_against_oracle(dmgx, cu256k, oracle_cu256k):
    s, ref = cu256k, oracle_cu256k
    N = len(s["types"])
    assert N == 256000
    out = gpu_compute(dmgx, dmgx.NEP(C1), s["box"], s["types"], s["pos"])
    err = errors(out, ref)
    assert err["force"] <= 1e-4, err
    assert err["energy_per_atom"] <= 1e-6, err
    assert err["energy_atom_max"] <= 2e-5, err
    assert err["virial"] <= 2e-4, err
    # momentum: the forces of a periodic system sum to zero
    assert np.abs(out[1].sum(1)).max() < 1e-6 * N


def test_fullsize_bf16_ann_error(dmgx, cu256k, oracle_cu256k):
    """BF16 operands on the matrix cores: the error against FP64 is that of
    bfloat16 rounding of the ANN inputs and weights, reported and bounded."""
    s, ref = cu256k, oracle_cu256k
    out = gpu_compute(dmgx, dmgx.NEP(C1), s["box"], s["types"], s["pos"], ann_bf16=True)
    err = errors(out, ref)
    print("bf16 ANN vs FP64:", err)
    assert err["force"] <= 0.05, err
    assert err["energy_per_atom"] <= 1e-3, err


def _nve_energies(dmgx, s, dt, chunk, chunks=20):
    N = len(s["types"])
    md = dmgx.MD(dmgx.NEP(C1), s["box"], s["types"], s["mass"], s["pos"], s["vel"])
    th = md.run(0, dt, "nve")
    e = [(1.5 * N * K_B * th[0] + th[1]) / N]
    for _ in range(chunks):
        th = md.run(chunk, dt, "nve")
        e.append((1.5 * N * K_B * th[0] + th[1]) / N)
    return md, np.array(e)


def test_fullsize_nve_energy_drift(dmgx, cu256k):
    """1000 fs of NVE from 300 K at dt 0.5 fs: total energy per atom within
    1e-5 eV of its start, total momentum unchanged.  The same 1000 fs at 1 fs
    moves it 4x further (1.6e-5 eV/atom): the excursion scales as dt^2, i.e.
    it is velocity Verlet's discretisation error, and the arithmetic (FP32
    forces, FP64 integration) adds less than the 1e-6 eV/atom seen at 0.25 fs."""
    s = cu256k
    N = len(s["types"])
    p0 = (s["mass"][None, :] * s["vel"]).sum(1)
    md, e_half = _nve_energies(dmgx, s, 0.5, 100)
    _, e_one = _nve_energies(dmgx, s, 1.0, 50)
    x_half = np.abs(e_half - e_half[0]).max()
    x_one = np.abs(e_one - e_one[0]).max()
    slope = np.polyfit(np.arange(len(e_half)) * 50.0, e_half, 1)[0]
    print(f"NVE 256k Cu over 1000 fs: max excursion {x_half:.3e} eV/atom at 0.5 fs, {x_one:.3e} at 1 fs; "
          f"slope {slope:.3e} eV/atom/fs at 0.5 fs")
    assert x_half <= 1e-5, (x_half, e_half)
    assert 3.5 <= x_one / x_half <= 4.5, (x_one, x_half)
    v = md.arrays()["velocity"]
    p1 = (s["mass"][None, :] * v).sum(1)
    assert np.abs(p1 - p0).max() / N < 1e-10, (p0, p1)
