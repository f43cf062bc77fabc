"""Parts of the drop-in contract beyond the plain evaluation, through the C
ABI: Verlet rows keyed by global id as the device stores them (the
reference's domain row order, neighbor.cuh:164-196, and its order test
tests/gpu/domain_neighbor_cuda_tests.cu:78-83); a domain share whose last
slots are coordinate-only ghosts (centres [0, n2) and [0, nd2) with
nd2 < N, nep.cu:1282-1292); and Nose-Hoover-chain MD against an FP64 replay
of the reference's ensemble step (ensemble_control.cpp:180-218,
runtime.cu:765-782)."""
import ctypes
import os

import numpy as np
import pytest

from conftest import POTENTIALS
from gpu_helpers import errors, gpu_compute, to_dev
from oracle import OracleNEP, neighbor_rows, thermo, velocity_verlet

pytestmark = pytest.mark.gpu

CARBON = os.path.join(POTENTIALS, "nep_carbon.txt")
CARBON_ZBL = os.path.join(POTENTIALS, "nep_carbon_zbl.txt")
TOL_F, TOL_E, TOL_EA, TOL_V = 1e-4, 1e-6, 2e-5, 2e-4
TIME_UNIT = 10.18051
KB = 8.617343e-5


def test_rows_keyed_by_global_id_in_device_order(dmgx, workloads):
    """Under DMGX_SORT_BY_GID every Verlet row, read back in the order the
    device stores it, is ascending in global id, holds exactly the
    brute-force neighbour set, and the forces are unchanged."""
    import torch
    s = workloads.carbon_diamond_4096(seed=2)
    N = len(s["types"])
    gid = np.random.default_rng(8).permutation(N).astype(np.int64) * 7 + 3
    nep = dmgx.NEP(CARBON)
    out = gpu_compute(dmgx, nep, s["box"], s["types"], s["pos"], global_id=to_dev(gid, torch.int64),
                      sort_by_gid=True)
    rows = nep.verlet_rows(0, N)
    ref = neighbor_rows(s["box"].h, s["box"].pbc, s["pos"], nep.info.rc_radial_max + 1.0, max_nb=1024)
    unsorted = 0
    for i in range(N):
        g = gid[rows[i]]
        assert np.all(np.diff(g) > 0), i
        assert np.array_equal(np.sort(rows[i]), ref[i]), i
        unsorted += int(np.any(np.diff(rows[i]) < 0))
    assert unsorted > N // 2  # the gid order is not the slot order
    orc = OracleNEP(CARBON, round_params_to_float=True).compute(s["box"].h, s["box"].pbc, s["types"], s["pos"])
    err = errors(out, orc)
    assert err["force"] <= TOL_F and err["virial"] <= TOL_V and err["energy_per_atom"] <= TOL_E, err


def _share_crop(workloads, core=7.0, rc=4.2, seed=5):
    """An open-box crop of an amorphous-carbon melt: `core` atoms (cube of
    half-width `core` around the centre) as force centres, every atom within
    rc + 0.5 of the core (cube metric) as further dependency centres, and
    the atoms within another rc + 0.5 as coordinate-only slots."""
    s = workloads.amorphous_carbon(6000, seed=seed)
    x = np.ascontiguousarray(s["pos"]).reshape(3, -1).T
    L = s["box"].h[0, 0]
    d = np.abs(x - L / 2).max(1)
    core_i = np.nonzero(d < core)[0]
    dep_i = np.nonzero((d >= core) & (d < core + rc + 0.5))[0]
    rest_i = np.nonzero((d >= core + rc + 0.5) & (d < core + 2.0 * rc + 1.0))[0]
    idx = np.concatenate([core_i, dep_i, rest_i])
    return x[idx], len(core_i), len(core_i) + len(dep_i), L


def test_coordinate_only_ghost_slots(dmgx, workloads):
    """Force centres [0, n2), dependency centres [0, nd2) and coordinate-only
    slots [nd2, N) of an open-box crop: owned forces, energies and virials
    against the oracle with the same centre definitions, and equal to the
    evaluation in which every slot is a dependency centre (the dependency
    shell covers every atom the force centres' partials reach)."""
    xyz, n2, nd2, L = _share_crop(workloads)
    N = xyz.shape[0]
    assert 0 < n2 < nd2 < N
    box = dmgx.Box.orthogonal(L, L, L, pbc=(0, 0, 0))
    types = np.zeros(N, np.int32)
    pos = np.ascontiguousarray(xyz.T)
    ref = OracleNEP(CARBON_ZBL, round_params_to_float=True).compute(box.h, box.pbc, types, pos, n1=0, n2=n2,
                                                                   nd1=0, nd2=nd2)
    part = gpu_compute(dmgx, dmgx.NEP(CARBON_ZBL), box, types, pos, n1=0, n2=n2, nd1=0, nd2=nd2)
    full = gpu_compute(dmgx, dmgx.NEP(CARBON_ZBL), box, types, pos, n1=0, n2=n2, nd1=0, nd2=N)
    err = errors(part, ref, n=n2)
    print(f"coordinate-only ghosts: {n2} force centres, {nd2 - n2} more dependency centres, "
          f"{N - nd2} coordinate-only slots; error {err}", flush=True)
    assert err["force"] <= TOL_F and err["virial"] <= TOL_V and err["energy_atom_max"] <= TOL_EA, err
    for a, b in zip(part, full):
        assert np.abs(a[..., :n2] - b[..., :n2]).max() < 1e-9


def test_nhc_run_matches_oracle_replay(dmgx, workloads):
    """Ten nvt_nhc steps of the 4,096-atom carbon system (300 -> 350 K ramp,
    Tdamp 20 steps, dt 1 fs) against an FP64 replay: oracle forces,
    integrator and thermo, with the library's bitwise chain (M = 4,
    thermostat_nhc.cpp:30-117) stepped as the reference's ensemble step
    does: a half step on the kinetic energy before the force, the other
    after it (ensemble_control.cpp:180-218, runtime.cu:765-782)."""
    s = workloads.carbon_diamond_4096(seed=6)
    N = len(s["types"])
    steps, dt_fs, T1, T2, tdamp = 10, 1.0, 300.0, 350.0, 20.0
    md = dmgx.MD(dmgx.NEP(CARBON), s["box"], s["types"], s["mass"], s["pos"], s["vel"])
    th_gpu = md.run(steps, dt_fs, "nvt_nhc", t_init=T1, t_final=T2, t_coupling=tdamp)
    a = md.arrays()
    L = dmgx._lib.lib()
    h = np.array(s["box"].h, float).reshape(3, 3)
    pos = np.ascontiguousarray(s["pos"], float).reshape(3, N).copy()
    vel = np.ascontiguousarray(s["vel"], float).reshape(3, N).copy()
    mass = np.asarray(s["mass"], float)
    orc = OracleNEP(CARBON, round_params_to_float=True)
    dt = dt_fs / TIME_UNIT
    chain = dmgx._lib.Nhc()
    assert L.dmgx_nhc_init(N, T1, tdamp, dt, ctypes.byref(chain)) == 0
    r = orc.compute(h, (1, 1, 1), s["types"], pos)
    f = np.ascontiguousarray(r["force"]).reshape(3, N)
    th = None
    for step in range(steps):
        target = T1 + (T2 - T1) * step / steps
        t0 = thermo(h, mass, r["pe"], vel, r["virial"])[0]
        ke2 = t0 * 3.0 * N * KB  # twice the kinetic energy
        vel *= L.dmgx_nhc_half_step(ctypes.byref(chain), ke2, KB * target, 3.0 * N, 0.5 * dt)
        velocity_verlet(1, dt, mass, pos, vel, f)
        r = orc.compute(h, (1, 1, 1), s["types"], pos)
        f = np.ascontiguousarray(r["force"]).reshape(3, N)
        velocity_verlet(0, dt, mass, pos, vel, f)
        th = thermo(h, mass, r["pe"], vel, r["virial"])
        vel *= L.dmgx_nhc_half_step(ctypes.byref(chain), th[0] * 3.0 * N * KB, KB * target, 3.0 * N, 0.5 * dt)
    Lb = h[0, 0]
    d = a["position"] - pos
    d -= Lb * np.round(d / Lb)
    dv = np.abs(a["velocity"] - vel).max()
    print(f"nvt_nhc 10 steps: position diff {np.abs(d).max():.2e} A, velocity diff {dv:.2e}, "
          f"T gpu {th_gpu[0]:.4f} replay {th[0]:.4f} K", flush=True)
    assert np.abs(d).max() < 1e-6
    assert dv < 1e-6 * max(1.0, np.abs(vel).max())
    # thermo of the last step, before its second chain half step
    assert abs(th_gpu[0] - th[0]) < 1e-6 * th[0]
    assert abs(th_gpu[1] - th[1]) < 1e-6 * N
