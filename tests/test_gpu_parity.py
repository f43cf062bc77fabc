"""Parity of the HIP NEP step with the FP64 oracle, through the C ABI.

Tolerances (stated against the FP64 oracle evaluated with the same
float32-stored parameters): force max |dF| <= 1e-4 eV/A, total energy
|dE|/N <= 1e-6 eV/atom, per-atom energy <= 2e-5 eV, per-atom virial
<= 2e-4 eV, all at the seeded configurations' real shapes; neighbour rows
bit-exact under the FP64 predicate."""
import os

import numpy as np
import pytest

from conftest import BASELINE, POTENTIALS
from gpu_helpers import errors, gpu_compute, to_dev
from oracle import OracleNEP, neighbor_rows
from xyz_io import XYZ_VIRIAL_TO_SOA, read_xyz, types_of

pytestmark = pytest.mark.gpu

C0 = os.path.join(POTENTIALS, "nep_c0_carbon.txt")
C1 = os.path.join(POTENTIALS, "nep_c1_cu.txt")
C4 = os.path.join(POTENTIALS, "nep_c4_carbon_zbl.txt")

TOL_F, TOL_E, TOL_EA, TOL_V = 1e-4, 1e-6, 2e-5, 2e-4


def check(err, scale_f=1.0, tol_e=TOL_E):
    assert err["force"] <= TOL_F * scale_f, err
    assert err["energy_per_atom"] <= tol_e, err
    assert err["energy_atom_max"] <= TOL_EA * scale_f, err
    assert err["virial"] <= TOL_V * scale_f, err


@pytest.mark.parametrize("case,stage,pot,fscale", [
    ("single_large_nve", "main", "nep_C.txt", 1.0),
    ("multi_nvt_restart", "initial", "nep_water.txt", 1.0),
    ("nep_zbl_boundary", "main", "nep_BaTiO3_zbl.txt", 50.0),  # 1.2e4 eV/A ZBL pair in FP32
])
def test_golden_systems_against_oracle_and_gpumd(dmgx, case, stage, pot, fscale):
    g = read_xyz(os.path.join(BASELINE, "goldens", case, stage, "initial.xyz"))
    path = os.path.join(BASELINE, "inputs", "potentials", pot)
    t = types_of(g["species"], path)
    pos = g["pos"].T.copy()
    box = dmgx.Box(g["h"], tuple(g["pbc"]))
    out = gpu_compute(dmgx, dmgx.NEP(path), box, t, pos)
    ref = OracleNEP(path, round_params_to_float=True).compute(g["h"], g["pbc"], t, pos)
    # 8/9-atom systems: the mean energy error cannot average below the FP32
    # per-atom floor (the reference's own FP32 outputs sit at 5e-6 eV/atom from
    # FP64), so the energy bound here is the reference manifest's 1e-5 relative.
    check(errors(out, ref), fscale, tol_e=1e-5 * np.abs(ref["pe"]).mean())
    # and against the reference engine's own FP32 outputs at its manifest tolerances
    pe, f, v = out
    assert np.all(np.abs(f.T - g["forces"]) <= 1e-6 + 1e-4 * np.abs(g["forces"]))
    assert np.all(np.abs(pe - g["energy_atom"]) <= 1e-8 + 1e-5 * np.abs(g["energy_atom"]))
    gv = g["virial"][:, XYZ_VIRIAL_TO_SOA]
    assert np.all(np.abs(v.T - gv) <= 1e-6 + 1e-4 * np.abs(gv))


def _system(workloads, which):
    if which == "c0":
        return workloads.carbon_diamond_4096(seed=1), C0
    if which == "c1":
        return workloads.copper_fcc(n=10, seed=3), C1
    if which == "c4":
        return workloads.carbon_diamond_4096(seed=5), C4
    raise KeyError(which)


@pytest.mark.parametrize("which", ["c0", "c1", "c4"])
def test_config_force_energy_virial(dmgx, workloads, which):
    s, path = _system(workloads, which)
    out = gpu_compute(dmgx, dmgx.NEP(path), s["box"], s["types"], s["pos"])
    ref = OracleNEP(path, round_params_to_float=True).compute(s["box"].h, s["box"].pbc, s["types"], s["pos"])
    check(errors(out, ref))
    assert abs(out[1].sum(1)).max() < 1e-6 * len(s["types"])  # momentum conservation


@pytest.mark.parametrize("which", ["c0", "c1"])
def test_neighbor_rows_bit_exact(dmgx, workloads, which):
    s, path = _system(workloads, which)
    nep = dmgx.NEP(path)
    gpu_compute(dmgx, nep, s["box"], s["types"], s["pos"])
    N = len(s["types"])
    info = nep.info
    rv = info.rc_radial_max + 1.0
    for rows_gpu, rc in ((nep.verlet_rows(0, N), rv), (nep.radial_rows(0, N), info.rc_radial_max),
                         (nep.angular_rows(0, N), info.rc_angular_max)):
        rows_ref = neighbor_rows(s["box"].h, s["box"].pbc, s["pos"], rc, max_nb=1024)
        assert len(rows_gpu) == N
        bad = [i for i in range(N) if not np.array_equal(rows_gpu[i], rows_ref[i])]
        assert not bad, (rc, bad[:5])


def test_deterministic_and_accumulate(dmgx, workloads):
    import torch
    s, path = _system(workloads, "c0")
    nep = dmgx.NEP(path)
    a = gpu_compute(dmgx, nep, s["box"], s["types"], s["pos"], force_rebuild=True)
    b = gpu_compute(dmgx, nep, s["box"], s["types"], s["pos"])
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    N = len(s["types"])
    t, x = to_dev(s["types"], torch.int32), to_dev(s["pos"], torch.float64)
    pe = torch.ones(N, dtype=torch.float64, device="cuda")
    f = torch.ones(3 * N, dtype=torch.float64, device="cuda")
    v = torch.ones(9 * N, dtype=torch.float64, device="cuda")
    nep.compute(s["box"], t, x, pe, f, v, accumulate=True)
    torch.cuda.synchronize()
    assert np.allclose(pe.cpu().numpy(), a[0] + 1.0, rtol=0, atol=1e-12)
    assert np.allclose(f.cpu().numpy(), a[1].reshape(-1) + 1.0, rtol=0, atol=1e-12)
    assert np.allclose(v.cpu().numpy(), a[2].reshape(-1) + 1.0, rtol=0, atol=1e-12)


def test_verlet_reuse_within_skin_and_rebuild(dmgx, workloads):
    s, path = _system(workloads, "c1")
    nep = dmgx.NEP(path)
    gpu_compute(dmgx, nep, s["box"], s["types"], s["pos"])
    r0 = nep.neighbor_stats()["rebuilds"]
    rng = np.random.default_rng(7)
    small = s["pos"] + rng.uniform(-0.15, 0.15, s["pos"].shape)  # < skin/2 = 0.5 A
    out = gpu_compute(dmgx, nep, s["box"], s["types"], small)
    assert nep.neighbor_stats()["rebuilds"] == r0
    ref = OracleNEP(path, round_params_to_float=True).compute(s["box"].h, s["box"].pbc, s["types"], small)
    check(errors(out, ref))
    big = s["pos"].copy()
    big[:, 0] += 0.6
    gpu_compute(dmgx, nep, s["box"], s["types"], big)
    assert nep.neighbor_stats()["rebuilds"] == r0 + 1


def test_domain_ranges(dmgx, workloads):
    """Force centres [0,n2) of a full-dependency evaluation equal the plain
    call; dependency centres restricted to what the force centres need."""
    s, path = _system(workloads, "c0")
    nep = dmgx.NEP(path)
    N = len(s["types"])
    full = gpu_compute(dmgx, nep, s["box"], s["types"], s["pos"])
    part = gpu_compute(dmgx, dmgx.NEP(path), s["box"], s["types"], s["pos"], n1=0, n2=1000, nd1=0, nd2=N)
    assert np.abs(part[1][:, :1000] - full[1][:, :1000]).max() < 1e-9
    assert np.abs(part[2][:, :1000] - full[2][:, :1000]).max() < 1e-9
    gid = np.arange(N, dtype=np.int64)[::-1].copy()
    import torch
    sorted_gid = gpu_compute(dmgx, dmgx.NEP(path), s["box"], s["types"], s["pos"],
                             global_id=to_dev(gid, torch.int64), sort_by_gid=True)
    ref = OracleNEP(path, round_params_to_float=True).compute(s["box"].h, s["box"].pbc, s["types"], s["pos"])
    check(errors(sorted_gid, ref))


def test_thermo_and_integrator_match_oracle(dmgx, workloads):
    import torch
    import oracle as O
    s, path = _system(workloads, "c0")
    N = len(s["types"])
    rng = np.random.default_rng(2)
    f = rng.normal(size=(3, N))
    vir = rng.normal(size=(9, N))
    pe = rng.normal(size=N)
    x = s["pos"].copy()
    v = s["vel"].copy()
    dx, dv, df = to_dev(x, torch.float64), to_dev(v, torch.float64), to_dev(f, torch.float64)
    dm = to_dev(s["mass"], torch.float64)
    L = dmgx.lib()
    from ctypes import c_void_p as P
    dt = 1.0 / 10.18051
    dmgx._lib.check(L.dmgx_velocity_verlet(1, N, dt, P(dm.data_ptr()), P(dx.data_ptr()), P(dv.data_ptr()),
                                           P(df.data_ptr()), None))
    O.velocity_verlet(True, dt, s["mass"], x, v, f)
    torch.cuda.synchronize()
    assert np.array_equal(dx.cpu().numpy(), x.reshape(-1))
    assert np.array_equal(dv.cpu().numpy(), v.reshape(-1))
    th = np.zeros(8)
    b = s["box"].to_c()
    import ctypes
    dpe, dvir = to_dev(pe, torch.float64), to_dev(vir, torch.float64)  # keep both alive
    dmgx._lib.check(L.dmgx_thermo(ctypes.byref(b), N, P(dm.data_ptr()), P(dpe.data_ptr()), P(dv.data_ptr()),
                                  P(dvir.data_ptr()), th.ctypes.data_as(dmgx._lib.PD), None))
    ref = O.thermo(s["box"].h, s["mass"], pe, v, vir)
    assert np.allclose(th, ref, rtol=1e-14, atol=1e-15)


def test_nve_trajectory_matches_oracle(dmgx, workloads):
    """20 NVE steps at 1 fs of the 4096-atom carbon config: positions,
    velocities and energies against an FP64 oracle trajectory."""
    import oracle as O
    s, path = _system(workloads, "c0")
    md = dmgx.MD(dmgx.NEP(path), s["box"], s["types"], s["mass"], s["pos"], s["vel"])
    steps, dt = 20, 1.0
    th = md.run(steps, dt, "nve")
    got = md.arrays()
    nep = OracleNEP(path, round_params_to_float=True)
    x, v = s["pos"].copy(), s["vel"].copy()
    h, pbc = s["box"].h, s["box"].pbc
    o = nep.compute(h, pbc, s["types"], x)
    dtn = dt / 10.18051
    for _ in range(steps):
        O.velocity_verlet(True, dtn, s["mass"], x, v, o["force"])
        O.wrap_positions(h, pbc, x)
        o = nep.compute(h, pbc, s["types"], x)
        O.velocity_verlet(False, dtn, s["mass"], x, v, o["force"])
    ref_th = O.thermo(h, s["mass"], o["pe"], v, o["virial"])
    d = got["position"] - x
    d -= np.round(d / 28.536) * 28.536
    assert np.abs(d).max() < 1e-6
    assert np.abs(got["velocity"] - v).max() < 1e-6
    assert abs(th[1] - ref_th[1]) / len(s["types"]) < 1e-6
    assert abs(th[0] - ref_th[0]) < 1e-3


def test_rejects_host_tensors_and_bad_ranges(dmgx, workloads):
    import torch
    s, path = _system(workloads, "c0")
    nep = dmgx.NEP(path)
    N = len(s["types"])
    t = torch.tensor(s["types"], dtype=torch.int32)
    with pytest.raises(ValueError):
        nep.compute(s["box"], t, torch.zeros(3 * N, dtype=torch.float64), torch.zeros(N, dtype=torch.float64),
                    torch.zeros(3 * N, dtype=torch.float64), torch.zeros(9 * N, dtype=torch.float64))
    with pytest.raises(dmgx.DmgxError):
        gpu_compute(dmgx, nep, s["box"], s["types"], s["pos"], n1=10, n2=20, nd1=15, nd2=N)
