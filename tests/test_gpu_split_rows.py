"""The typewise split of the Verlet rows done once per rebuild
(dmgx_nep_set_split(1): candidate rows of every Verlet entry within its
pair's rc + skin) instead of the per-step filter pass of the reference
(nep.cu:503-561), through the C ABI: the FP64-oracle bars of the plain
path, the same step's angular rows bit-exact, the per-step filter's forces
to FP32 summation order (the radial sums visit the same edges in another
order), and the same MD trajectories, Verlet reuse across box rescales and
serving batches as the per-step filter."""
import os

import numpy as np
import pytest

from conftest import POTENTIALS
from gpu_helpers import errors, gpu_compute, to_dev
from oracle import OracleNEP, neighbor_rows
from test_gpu_parity import CARBON, CU, CARBON_ZBL, WATER, _npt_replay, _system, check

pytestmark = pytest.mark.gpu

# forces of the two row schemes: the same edges, radial sums in another
# FP32 order
TOL_MODES_F, TOL_MODES_E = 2e-5, 2e-6


def _both(dmgx, path, box, types, pos, **kw):
    filt, split = dmgx.NEP(path), dmgx.NEP(path)
    filt.set_split(False)
    split.set_split(True)
    a = gpu_compute(dmgx, filt, box, types, pos, **kw)
    b = gpu_compute(dmgx, split, box, types, pos, **kw)
    assert split.tuning()["split"] == 1 and filt.tuning()["split"] == 0
    return a, b, filt, split


def _modes_agree(a, b, n=None, scale=1.0):
    sl = slice(0, n)
    df = float(np.abs(a[1][:, sl] - b[1][:, sl]).max())
    de = float(np.abs(a[0][sl] - b[0][sl]).max())
    print(f"split rows vs per-step filter: max |dF| {df:.2e} eV/A, max |dE_i| {de:.2e} eV", flush=True)
    assert df <= TOL_MODES_F * scale and de <= TOL_MODES_E * scale, (df, de)


@pytest.mark.parametrize("which", ["carbon", "cu", "carbon_zbl"])
def test_split_rows_against_oracle(dmgx, workloads, which):
    s, path = _system(workloads, which)
    a, b, _, split = _both(dmgx, path, s["box"], s["types"], s["pos"])
    ref = OracleNEP(path, round_params_to_float=True).compute(s["box"].h, s["box"].pbc, s["types"], s["pos"])
    check(errors(b, ref))
    _modes_agree(a, b)
    assert abs(b[1].sum(1)).max() < 1e-6 * len(s["types"])


@pytest.mark.parametrize("which", ["carbon", "cu"])
def test_split_rows_step_rows(dmgx, workloads, which):
    """The angular rows the descriptor writes from the candidate rows and the
    radial rows of the step are the brute-force FP64 rows, and the row
    totals equal the per-step filter's."""
    s, path = _system(workloads, which)
    _, _, filt, split = _both(dmgx, path, s["box"], s["types"], s["pos"])
    N = len(s["types"])
    info = split.info
    for rows_gpu, rc in ((split.angular_rows(0, N), info.rc_angular_max), (split.radial_rows(0, N), info.rc_radial_max)):
        rows_ref = neighbor_rows(s["box"].h, s["box"].pbc, s["pos"], rc, max_nb=1024)
        bad = [i for i in range(N) if not np.array_equal(rows_gpu[i], rows_ref[i])]
        assert not bad, (rc, bad[:5])
    ma, mb = filt.neighbor_means(), split.neighbor_means()
    assert ma == mb, (ma, mb)


@pytest.mark.parametrize("name", ["nep4_typecut_ch.txt", "nep5_zbl_typewise_ch.txt", "nep4_lmax6_ch.txt"])
def test_split_rows_per_type_cutoffs(dmgx, name):
    """Per-type-pair cutoffs (the candidate rows drop Verlet entries beyond
    their pair's rc + skin), the typewise ZBL and the order-8 instance on the
    3x3x3 tiling of the random 2-type (C, H) cell, against the oracle with
    the bars of the per-step path."""
    from test_oracle_golden import coverage_sample
    h, pos, t = coverage_sample()
    L = h[0, 0]
    shifts = np.array([[a, b, c] for a in range(3) for b in range(3) for c in range(3)]) * L
    pos = np.ascontiguousarray(np.concatenate([pos + s[:, None] for s in shifts], axis=1))
    t = np.tile(t, 27)
    h = h * 3
    path = os.path.join(POTENTIALS, name)
    a, b, _, _ = _both(dmgx, path, dmgx.Box(h, (1, 1, 1)), t, pos)
    ref = OracleNEP(path, round_params_to_float=True).compute(h, (1, 1, 1), t, pos)
    err = errors(b, ref)
    fmax, emax = np.abs(ref["force"]).max(), np.abs(ref["pe"]).max()
    assert err["force"] <= max(1e-4, 2e-6 * fmax), (err, fmax)
    assert err["energy_atom_max"] <= max(2e-5, 2e-7 * emax), (err, emax)
    assert err["virial"] <= max(2e-4, 2e-6 * np.abs(ref["virial"]).max()), err
    _modes_agree(a, b, scale=max(1.0, fmax / 10.0))


def test_split_rows_water(dmgx, workloads):
    s = workloads.liquid_water(1000, seed=3)
    a, b, _, _ = _both(dmgx, WATER, s["box"], s["types"], s["pos"])
    ref = OracleNEP(WATER, round_params_to_float=True).compute(s["box"].h, s["box"].pbc, s["types"], s["pos"])
    fscale = max(1.0, np.abs(ref["force"]).max() / 10.0)
    check(errors(b, ref), scale_f=fscale)
    _modes_agree(a, b, scale=fscale)


def test_split_rows_coordinate_only_ghosts(dmgx, workloads):
    """Force centres [0, n2), dependency centres [0, nd2) and coordinate-only
    slots of an open-box crop (rows only for dependency centres, the reverse
    partials only from them), against the oracle with the same centres."""
    from test_gpu_contract import _share_crop
    xyz, n2, nd2, L = _share_crop(workloads)
    N = xyz.shape[0]
    box = dmgx.Box.orthogonal(L, L, L, pbc=(0, 0, 0))
    types = np.zeros(N, np.int32)
    pos = np.ascontiguousarray(xyz.T)
    ref = OracleNEP(CARBON_ZBL, round_params_to_float=True).compute(box.h, box.pbc, types, pos, n1=0, n2=n2,
                                                                   nd1=0, nd2=nd2)
    a, b, _, _ = _both(dmgx, CARBON_ZBL, box, types, pos, n1=0, n2=n2, nd1=0, nd2=nd2)
    err = errors(b, ref, n=n2)
    assert err["force"] <= 1e-4 and err["virial"] <= 2e-4 and err["energy_atom_max"] <= 2e-5, err
    _modes_agree(a, b, n=n2)


def test_split_rows_reuse_and_rebuild(dmgx, workloads):
    """Within the skin the candidate rows are reused (no rebuild) and the
    step still matches the oracle; a displacement past half the skin
    rebuilds them."""
    s, path = _system(workloads, "cu")
    nep = dmgx.NEP(path)
    nep.set_split(True)
    gpu_compute(dmgx, nep, s["box"], s["types"], s["pos"])
    r0 = nep.neighbor_stats()["rebuilds"]
    rng = np.random.default_rng(7)
    small = s["pos"] + rng.uniform(-0.15, 0.15, s["pos"].shape)
    out = gpu_compute(dmgx, nep, s["box"], s["types"], small)
    assert nep.neighbor_stats()["rebuilds"] == r0
    ref = OracleNEP(path, round_params_to_float=True).compute(s["box"].h, s["box"].pbc, s["types"], small)
    check(errors(out, ref))
    big = s["pos"].copy()
    big[:, 0] += 0.6
    out = gpu_compute(dmgx, nep, s["box"], s["types"], big)
    assert nep.neighbor_stats()["rebuilds"] == r0 + 1
    ref = OracleNEP(path, round_params_to_float=True).compute(s["box"].h, s["box"].pbc, s["types"], big)
    check(errors(out, ref))


def test_split_rows_nve_matches_filter(dmgx, workloads):
    """Forty NVE steps of the 4,096-atom carbon system started at 3,000 K
    (six rebuilds on the device) with candidate rows and with the per-step
    filter: the same rebuild decisions, and trajectories that agree to the
    growth of the FP32 summation-order difference."""
    s = workloads.carbon_diamond_4096(seed=11)
    vel = np.ascontiguousarray(s["vel"]) * np.sqrt(10.0)
    out = []
    for split in (False, True):
        nep = dmgx.NEP(CARBON)
        nep.set_split(split)
        md = dmgx.MD(nep, s["box"], s["types"], s["mass"], s["pos"], vel)
        th = md.run(40, 1.0, "nve")
        out.append((md.arrays(), th, nep.neighbor_stats()["rebuilds"]))
    (a, tha, ra), (b, thb, rb) = out
    L = s["box"].h[0, 0]
    d = a["position"] - b["position"]
    d -= L * np.round(d / L)
    print(f"split vs filter NVE: position diff {np.abs(d).max():.2e} A, rebuilds {ra} / {rb}, "
          f"T {tha[0]:.4f} / {thb[0]:.4f} K", flush=True)
    # the system heats to about 4,300 K: the 1e-6 eV/A force differences of
    # the two summation orders grow to about 1e-5 A over the forty steps
    assert ra == rb
    assert np.abs(d).max() < 5e-5
    assert abs(tha[0] - thb[0]) < 1e-5 * tha[0]


def test_split_rows_npt_matches_oracle_replay(dmgx, workloads):
    """Ten Berendsen NPT steps of 1000 water molecules (tau_p 1000 steps:
    the rows are kept across the box rescales) with candidate rows, against
    the FP64 replay of the reference's ensemble step."""
    s = workloads.liquid_water(1000, seed=4)
    kw = dict(steps=10, dt_fs=0.5, T=300.0, P=1e-4, B=2.2, tau_p=1000.0, tau_t=100.0, seed=12345)
    nep = dmgx.NEP(WATER)
    nep.set_split(True)
    md = dmgx.MD(nep, s["box"], s["types"], s["mass"], s["pos"], s["vel"])
    md.run(kw["steps"], kw["dt_fs"], "npt_ber", t_init=kw["T"], t_final=kw["T"], t_coupling=kw["tau_t"],
           p_target_gpa=kw["P"], bulk_modulus_gpa=kw["B"], p_coupling=kw["tau_p"], seed=kw["seed"])
    a = md.arrays()
    h_gpu = np.array(md.box().h, float).reshape(3, 3)
    h, pos, vel = _npt_replay(dmgx, s, "npt_ber", **kw)
    Lr = h[0, 0]
    d = a["position"] - pos
    d -= Lr * np.round(d / Lr)
    rebuilds = nep.neighbor_stats()["rebuilds"]
    print(f"split rows npt_ber: box diff {abs(h_gpu[0, 0] - Lr):.2e}, pos diff {np.abs(d).max():.2e}, "
          f"rebuilds {rebuilds}", flush=True)
    assert rebuilds < kw["steps"]
    assert abs(h_gpu[0, 0] / Lr - 1.0) < 1e-8
    assert np.abs(d).max() < 1e-6
    assert np.abs(a["velocity"] - vel).max() < 1e-5 * max(1.0, np.abs(vel).max())


def test_split_rows_serving_batch(dmgx):
    """A serving batch (owner-mapped image ghosts) on a handle with candidate
    rows: every request against the FP64 oracle of its periodic system."""
    from test_gpu_serving import _check_single, _diamond
    orc = OracleNEP(CARBON, round_params_to_float=True)
    nep = dmgx.NEP(CARBON)
    nep.set_split(True)
    srv = dmgx.Server(nep, max_atoms=400_000)
    cases = []
    for k in range(12):
        h, p = _diamond(1 + k % 3, 0.05, 300 + k)
        n = p.shape[1]
        cases.append((srv.submit(dmgx.Box(h, (1, 1, 1)), np.zeros(n, np.int32), np.full(n, 12.011), p), h, p))
    srv.drain()
    for rid, h, p in cases:
        _check_single(srv.poll(rid), h, p, orc)
    assert nep.tuning()["split"] == 1
