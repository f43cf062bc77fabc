"""The benchmark configuration at its full size: 256,000-atom Cu fcc with the
Cu NEP4 model.  Forces, energies and virials of every atom against the FP64
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

CU = os.path.join(POTENTIALS, "nep_cu.txt")
K_B = 8.617343e-5


@pytest.fixture(scope="module")
def cu256k(workloads):
    return workloads.copper_fcc(n=40, seed=1)


@pytest.fixture(scope="module")
def oracle_cu256k(cu256k):
    s = cu256k
    print("Cu 256k: FP64 oracle on every atom", flush=True)
    return OracleNEP(CU, round_params_to_float=True).compute(s["box"].h, s["box"].pbc, s["types"], s["pos"])


def test_fullsize_against_oracle(dmgx, cu256k, oracle_cu256k):
    s, ref = cu256k, oracle_cu256k
    N = len(s["types"])
    assert N == 256000
    out = gpu_compute(dmgx, dmgx.NEP(CU), s["box"], s["types"], s["pos"])
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
    out = gpu_compute(dmgx, dmgx.NEP(CU), s["box"], s["types"], s["pos"], ann_bf16=True)
    err = errors(out, ref)
    print("bf16 ANN vs FP64:", err)
    assert err["force"] <= 0.05, err
    assert err["energy_per_atom"] <= 1e-3, err


def test_fullsize_split_bf16_ann_meets_the_fp32_bars(dmgx, cu256k, oracle_cu256k):
    """Split bfloat16 operands on the matrix cores (every operand as hi + lo
    bfloat16 parts, three products per term, FP32 accumulation): every atom
    within the FP32 bars (1e-4 eV/A, 1e-6 eV/atom)."""
    s, ref = cu256k, oracle_cu256k
    out = gpu_compute(dmgx, dmgx.NEP(CU), s["box"], s["types"], s["pos"], ann_bf16_split=True)
    err = errors(out, ref)
    print("split-bf16 ANN vs FP64:", err)
    assert err["force"] <= 1e-4, err
    assert err["energy_per_atom"] <= 1e-6, err
    assert err["energy_atom_max"] <= 2e-5, err
    assert err["virial"] <= 2e-4, err


def _nve_energies(dmgx, s, dt, chunk, chunks=20):
    N = len(s["types"])
    md = dmgx.MD(dmgx.NEP(CU), s["box"], s["types"], s["mass"], s["pos"], s["vel"])
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


# ---- liquid water: 64,000 H2O (192,000 atoms), 2-type NEP4 ----------------

WATER = os.path.join(POTENTIALS, "nep_water.txt")


@pytest.fixture(scope="module")
def water192k(workloads):
    return workloads.liquid_water(64000, seed=1)


def test_water_fullsize_against_oracle(dmgx, water192k):
    """Every atom of the liquid-water configuration against the FP64 oracle;
    the force bar scales with the largest force (close intermolecular pairs
    of the random packing give forces of tens of eV/A)."""
    s = water192k
    N = len(s["types"])
    assert N == 192000
    print("water 192k: FP64 oracle on every atom", flush=True)
    ref = OracleNEP(WATER, round_params_to_float=True).compute(s["box"].h, s["box"].pbc, s["types"], s["pos"])
    out = gpu_compute(dmgx, dmgx.NEP(WATER), s["box"], s["types"], s["pos"])
    err = errors(out, ref)
    fmax = np.abs(ref["force"]).max()
    print("water 192k vs FP64:", err, "max |F|", fmax)
    assert err["force"] <= max(1e-4, 2e-6 * fmax), err
    assert err["energy_per_atom"] <= 1e-6, err
    assert err["energy_atom_max"] <= 2e-5, err
    assert err["virial"] <= max(2e-4, 2e-6 * np.abs(ref["virial"]).max()), err
    assert np.abs(out[1].sum(1)).max() < 1e-6 * N


def test_water_fullsize_scr_npt_run(dmgx, water192k):
    """The configuration's run: 500 stochastic-cell-rescaling NPT steps at
    300 K and 1e-4 GPa, dt 0.5 fs.  Its first two steps at full size against
    the FP64 replay (box, positions, velocities); the whole run finite with
    the box cubic.  (The random packing of a random-weight model is far from
    equilibrium: the barostat shrinks the box by ~14% over the run.)"""
    import time
    import torch
    from test_gpu_parity import _npt_replay
    s = water192k
    N = len(s["types"])
    kw = dict(T=300.0, P=1e-4, B=2.2, tau_p=1000.0, tau_t=100.0, seed=12345678)
    md = dmgx.MD(dmgx.NEP(WATER), s["box"], s["types"], s["mass"], s["pos"], s["vel"])
    run = lambda k: md.run(k, 0.5, "npt_scr", t_init=kw["T"], t_final=kw["T"], t_coupling=kw["tau_t"],
                           p_target_gpa=kw["P"], bulk_modulus_gpa=kw["B"], p_coupling=kw["tau_p"], seed=kw["seed"])
    run(2)
    a = md.arrays()
    hg = np.array(md.box().h, float).reshape(3, 3)
    h, pos, vel = _npt_replay(dmgx, s, "npt_scr", steps=2, dt_fs=0.5, model=WATER, **kw)
    L = h[0, 0]
    d = a["position"] - pos
    d -= L * np.round(d / L)
    print("water 192k SCR NPT, 2 steps vs FP64 replay: box", abs(hg[0, 0] / L - 1.0), "pos", np.abs(d).max(),
          "vel", np.abs(a["velocity"] - vel).max())
    assert abs(hg[0, 0] / L - 1.0) < 1e-9
    assert np.abs(d).max() < 1e-6
    assert np.abs(a["velocity"] - vel).max() < 1e-5 * max(1.0, np.abs(vel).max())
    md = dmgx.MD(dmgx.NEP(WATER), s["box"], s["types"], s["mass"], s["pos"], s["vel"])
    run(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    th = run(500)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    h0 = np.array(s["box"].h, float).reshape(3, 3)
    h = np.array(md.box().h, float).reshape(3, 3)
    a = md.arrays()
    print(f"water 192k SCR NPT 500 steps: {500 * N / el:.3e} atom-steps/s, T {th[0]:.1f} K, "
          f"volume change {np.linalg.det(h) / np.linalg.det(h0) - 1.0:.3e}")
    assert np.all(np.isfinite(th)) and all(np.all(np.isfinite(a[k])) for k in ("position", "velocity", "force"))
    assert np.allclose(h, np.diag(np.diag(h))) and np.ptp(np.diag(h)) < 1e-9 * h[0, 0]


# ---- one GPU's share: amorphous carbon, 2,048,000 atoms + 5 A ghost shell ----

CARBON_ZBL = os.path.join(POTENTIALS, "nep_carbon_zbl.txt")


@pytest.fixture(scope="module")
def carbon_share(workloads):
    print("a-C share: generating 2,048,000 atoms and the ghost shell", flush=True)
    return workloads.amorphous_carbon_share(2048000, seed=1)


def _crop_oracle(s, centre, core=10.0, rc=4.2):
    """Oracle forces of the owned atoms in the cube of half-width `core`
    around `centre`: they depend only on atoms within two cutoffs, so the
    oracle runs on that crop of the share (same open box) with the core as
    force centres, every atom within a cutoff of it as dependency centres
    (a superset, by the cube metric) and the rest of the crop as candidates,
    as the full-share evaluation defines them.  Returns (core indices,
    oracle result)."""
    x = np.ascontiguousarray(s["pos"]).reshape(3, -1).T
    d = np.abs(x - np.asarray(centre)).max(1)
    core_i = np.nonzero(d < core)[0]
    assert core_i.max() < s["owned"]  # the core holds owned atoms only
    dep_i = np.nonzero((d >= core) & (d < core + rc + 0.5))[0]
    rest_i = np.nonzero((d >= core + rc + 0.5) & (d < core + 2.0 * rc + 1.0))[0]
    idx = np.concatenate([core_i, dep_i, rest_i])
    sub = np.ascontiguousarray(x[idx].T)
    ref = OracleNEP(CARBON_ZBL, round_params_to_float=True).compute(s["box"].h, s["box"].pbc, s["types"][idx], sub,
                                                            n1=0, n2=len(core_i), nd1=0,
                                                            nd2=len(core_i) + len(dep_i))
    return core_i, ref


def _crop_centres(s):
    """Interior, against an owned face (its dependencies reach into the ghost
    shell) and at an owned corner."""
    o, L = s["origin"], s["L"]
    return [(o + L / 2,) * 3, (o + 10.0, o + L / 2, o + L / 2), (o + L - 10.0,) * 3]


def _share_compute(dmgx, nep, s, pos=None, bf16=False):
    import torch
    N, n = len(s["types"]), s["owned"]
    t = torch.tensor(s["types"], dtype=torch.int32, device="cuda")
    x = torch.tensor(np.ascontiguousarray(s["pos"]).reshape(-1), dtype=torch.float64, device="cuda") \
        if pos is None else pos
    pe = torch.zeros(N, dtype=torch.float64, device="cuda")
    f = torch.zeros(3 * N, dtype=torch.float64, device="cuda")
    v = torch.zeros(9 * N, dtype=torch.float64, device="cuda")
    nep.compute_domain(s["box"], N, t, x, pe, f, v, n1=0, n2=n, nd1=0, nd2=N, ann_bf16=bf16)
    return pe, f.view(3, N), v.view(9, N)


def test_carbon_share_against_oracle(dmgx, carbon_share):
    """The 2.28M-slot share through the domain contract (force centres = the
    owned atoms, dependency centres = owned + ghosts), FP32 and BF16 ANN:
    owned forces, energies and virials against the oracle on three crops
    with the same centre definitions.  Random positions 1.2 A apart put
    ZBL-range pairs everywhere.  FP32: every crop atom within 1e-4 eV/A,
    2e-4 eV per virial component and 2e-5 eV per atom energy (the model's
    scaler is fit on a melt sample; a host FP32 emulation of this melt
    puts the largest force error near 1.6e-5).  BF16 ANN operands: the
    error is that of bfloat16 rounding of the ANN inputs, reported and
    bounded relative to the largest force."""
    s = carbon_share
    n = s["owned"]
    assert n == 2048000
    crops = []
    for c in _crop_centres(s):
        print("a-C share: oracle on the crop around", tuple(round(x, 1) for x in c), flush=True)
        crops.append(_crop_oracle(s, c))
    nep = dmgx.NEP(CARBON_ZBL)
    for bf16 in (False, True):
        print("a-C share: device evaluation", "bf16" if bf16 else "fp32", flush=True)
        pe, f, v = _share_compute(dmgx, nep, s, bf16=bf16)
        pe, f, v = pe.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy()
        for core_i, ref in crops:
            m = len(core_i)
            err = errors((pe[core_i], f[:, core_i], v[:, core_i]), ref, n=m)
            fmax = np.abs(ref["force"][:, :m]).max()
            ferr = np.abs(f[:, core_i] - ref["force"][:, :m]).max(0)
            k = int(np.argmax(ferr))
            x = np.ascontiguousarray(s["pos"]).reshape(3, -1)
            dk = np.sort(np.linalg.norm(x[:, :s["owned"]] - x[:, [core_i[k]]], axis=0))[1]
            print("  worst atom |F|", float(np.abs(ref["force"][:, k]).max()), "err", float(ferr[k]),
                  "nearest", float(dk), "errors > 1e-4:", int((ferr > 1e-4).sum()), "of", m)
            print("a-C share", "bf16" if bf16 else "fp32", m, "atoms", err, "max |F|", fmax)
            if bf16:
                assert err["force"] <= 5e-2 * max(fmax, 1.0), err
                assert err["energy_atom_max"] <= 2e-2 * max(np.abs(ref["pe"][:m]).max(), 1.0), err
            else:
                assert err["force"] <= 1e-4, err
                assert err["energy_atom_max"] <= 2e-5, err
                assert err["virial"] <= 2e-4, err


def _share_nve(dmgx, s, dt_fs, steps, bf16):
    """NVE of the owned atoms with the ghost shell static; returns the
    conserved quantity per owned atom at every step: owned kinetic energy
    plus the energy of every dependency centre (ghosts included: their
    energies change as owned atoms move)."""
    import torch
    N, n = len(s["types"]), s["owned"]
    dt = dt_fs / 10.18051
    m = torch.tensor(s["mass"][:n], dtype=torch.float64, device="cuda")
    nep = dmgx.NEP(CARBON_ZBL)
    x = torch.tensor(np.ascontiguousarray(s["pos"]).reshape(-1), dtype=torch.float64, device="cuda")
    xv = x.view(3, N)
    v = torch.tensor(np.ascontiguousarray(s["vel"]), dtype=torch.float64, device="cuda").view(3, n)
    pe, f, _ = _share_compute(dmgx, nep, s, pos=x, bf16=bf16)
    h = []
    for step in range(steps + 1):
        h.append(float(0.5 * (m * (v * v).sum(0)).sum() + pe.sum()) / n)
        if step == steps:
            break
        v += 0.5 * dt * f[:, :n] / m
        xv[:, :n] += dt * v
        pe, f, _ = _share_compute(dmgx, nep, s, pos=x, bf16=bf16)
        v += 0.5 * dt * f[:, :n] / m
    return np.array(h)


def test_carbon_share_nve_with_static_ghosts(dmgx, carbon_share):
    """The configuration's run: 200 NVE steps at dt 0.5 fs of the owned atoms
    from 5000 K, the ghost shell static, FP32 and BF16 ANN.  The start is
    violent (ZBL pairs at 1.2 A push atoms apart at ~0.1 A/fs), so the
    conserved quantity moves by ~2e-2 eV/atom at 0.5 fs; the same 100 fs at
    0.25 fs moves it at least 3.5x less, i.e. the drift is velocity Verlet's
    discretisation error, not the forces'."""
    s = carbon_share
    h_half = _share_nve(dmgx, s, 0.5, 200, False)
    h_quarter = _share_nve(dmgx, s, 0.25, 400, False)
    h_bf16 = _share_nve(dmgx, s, 0.5, 200, True)
    d = {k: float(np.abs(h - h[0]).max()) for k, h in (("fp32 0.5 fs", h_half), ("fp32 0.25 fs", h_quarter),
                                                       ("bf16 0.5 fs", h_bf16))}
    print("a-C share NVE over 100 fs, max |dH| eV/atom:", d)
    assert all(np.isfinite(h).all() for h in (h_half, h_quarter, h_bf16))
    assert d["fp32 0.5 fs"] / d["fp32 0.25 fs"] >= 3.5, d
    assert d["fp32 0.5 fs"] < 0.05 and d["bf16 0.5 fs"] < 0.1, d
