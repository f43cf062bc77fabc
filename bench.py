"""Benchmark of the single-GPU NEP MD step: atom-steps/s on the 256,000-atom
Cu fcc NEP4 configuration (Bussi NVT, 300 K, dt 2 fs).

A "step" is one velocity-Verlet MD step: integrate, Verlet-list check (and
rebuild when an atom moved more than half the skin), typewise filter, fused
descriptor/ANN/angular-partial kernel, force/virial gather, integrate,
compensated thermo and the Bussi velocity rescale.  Inputs are resident in
HBM before the timed region.  Prints one JSON line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s (1 GPU, NEP4 Cu 256k); served frames/s + p99 latency at fixed force err"
MODEL = os.path.join(ROOT, "tests", "golden", "potentials", "nep_cu.txt")
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 (vector and MFMA), AMD public figure
PEAK_FP32_SCALAR_TFLOPS = 78.6  # the same without packed FP32 (v_pk_*_f32), which this build keeps off
PEAK_HBM_GBS = 8000.0


def flops_per_atom(info, n_r, n_a):
    """Algorithmic FLOP per atom-step of the reference algorithm (the
    canonical per-unit formulas of the measurement plan), split by kernel."""
    KR, KA = info.basis_size_radial + 1, info.basis_size_angular + 1
    NR, NA = info.n_max_radial + 1, info.n_max_angular + 1
    L = info.l_max
    ABC = (L + 1) ** 2 - 1
    D, H = info.dim, info.num_neurons
    q45 = info.num_L - L
    f_desc = n_r * (3 * KR + 2 * NR * KR + 10) + n_a * (3 * KA + 2 * NA * KA + 2 * NA * ABC + 60) + NA * (3 * ABC + 30)
    f_ann = 4 * D * H + 2 * H
    f_ang = n_a * (6 * KA + 4 * NA * KA + NA * (20 * ABC + 60 * q45) + 20)
    f_rad = n_r * (6 * KR + 4 * NR * KR + 6 * NR + 30)
    f_gather = 30 * n_a
    return {"fused": f_desc + f_ann + f_ang, "gather": f_rad + f_gather, "desc": f_desc, "ann": f_ann,
            "ang": f_ang, "rad": f_rad}


def measured_traffic(kernel_prefix: str, atoms: int):
    """HBM bytes per dispatch of the kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), when it covers this
    workload; None otherwise."""
    path = os.path.join(ROOT, "profiles", "cu256k_hbm_traffic.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None
    if d.get("atoms") != atoms:
        return None
    for name, v in d["kernels"].items():
        if name.startswith(kernel_prefix):
            return v["read_bytes"] + v["write_bytes"]
    return None


def cpu_baseline(budget_s=10.0):
    """The reference's own FP64 CPU evaluator (oracle/_ref, nep_cpu.cpp) when
    it was built, else the FP64 oracle port, timed on one host core on a
    6,912-atom sample of the same Cu configuration (forces + velocity Verlet)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    W = importlib.import_module("dmg-md_amd.workloads")
    s = W.copper_fcc(n=12, seed=7)
    kind = "port"
    try:
        import ref
        if ref.available():
            ev = ref.RefNEP(MODEL)
            kind = "reference"
        else:
            ev = O.OracleNEP(MODEL)
    except Exception:
        ev = O.OracleNEP(MODEL)
    x, v, m = s["pos"].copy(), s["vel"].copy(), s["mass"]
    h, pbc = s["box"].h, s["box"].pbc
    dt = 2.0 / 10.18051
    f = ev.compute(h, pbc, s["types"], x)["force"]
    steps = 0
    t0 = time.perf_counter()
    while True:
        O.velocity_verlet(True, dt, m, x, v, np.ascontiguousarray(f))
        O.wrap_positions(h, pbc, x)
        f = ev.compute(h, pbc, s["types"], x)["force"]
        O.velocity_verlet(False, dt, m, x, v, np.ascontiguousarray(f))
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 200:
            break
    n = len(s["types"])
    return {"value": n * steps / el, "unit": "atom-steps/s", "cores": 1, "kind": kind,
            "sample": f"Cu fcc 12^3 = {n} atoms, {steps} NVE steps (force + velocity Verlet), "
                      f"{el:.1f} s on 1 host core, FP64"}


CARBON_MODEL = os.path.join(ROOT, "tests", "golden", "potentials", "nep_carbon.txt")


def serve_bench(dmgx, requests=4096, rate_hz=1000.0, seed=7, capacity_only=False):
    """The serving configuration: `requests` carbon and Cu requests (each
    with its own model: the 4096-atom carbon and the 256k Cu configurations'
    NEP4 models, one server), 64..20000 sites log-uniform, 30% vacancies on a
    random subset, half single points (energy, forces, virial) and half
    50-step NVE segments.  The served mode is the FP32 ANN, the force error
    the plain path has; the bfloat16-operand ANN is measured beside it with
    its force error against FP32.
      capacity: the whole stream submitted at once by 8 client threads
        (closed loop, like the plain loop), frames/s from first submit to
        last completion, on a server warmed by one untimed pass of the same
        stream (its buffers, pools and the engines' workspaces sized);
      latency: the same stream with Poisson arrivals at rate_hz on the same
        server, p50/p99 submit-to-result;
      plain FP32: the same requests one at a time through NEP.compute /
        MD.run (FP32 ANN), after one untimed request per model."""
    import gc
    import numpy as np
    import torch
    from concurrent.futures import ThreadPoolExecutor
    W = importlib.import_module("dmg-md_amd.workloads")
    # device memory the earlier samples left in torch's cache goes back to
    # the device (the server's budget is half the free memory)
    gc.collect()
    torch.cuda.empty_cache()
    t_gen = time.perf_counter()
    stream = W.serving_stream(requests, seed)
    print(f"bench: serving stream of {requests} requests generated in {time.perf_counter() - t_gen:.1f} s",
          file=sys.stderr, flush=True)
    paths = [CARBON_MODEL, MODEL]
    gaps = np.random.default_rng(seed + 1).exponential(1.0 / rate_hz, requests)

    def submit(srv, k):
        r, model, steps = stream[k]
        return srv.submit(r["box"], r["types"], r["mass"], r["pos"], r["vel"], steps=steps, dt_fs=r["dt_fs"],
                          model=model)

    def burst(srv, count=requests, clients=8):
        srv.reset_stats()
        with ThreadPoolExecutor(clients) as pool:
            ids = list(pool.map(lambda k: submit(srv, k), range(count)))
        srv.drain()
        return ids, srv.stats()

    def poisson(srv):
        srv.reset_stats()
        t_next = time.perf_counter()
        for k, g in enumerate(gaps):
            t_next += g
            while time.perf_counter() < t_next:
                time.sleep(min(1e-4, max(0.0, t_next - time.perf_counter())))
            submit(srv, k)
        srv.drain()
        return srv.stats()

    def warmed(bf16):
        # the whole stream once, untimed: the server's layout buffers, pools
        # and the engines' workspaces reach their sizes, and each model's
        # autotuner picks its tile shapes on its first batch of >= 64k atoms.
        # The slot cap is set above what the budget admits: requests are
        # admitted by their HBM bytes (half the free device memory).
        srv = dmgx.Server(neps, max_atoms=40_000_000, ann_bf16=bf16)
        burst(srv)
        return srv

    neps = [dmgx.NEP(p) for p in paths]
    srv = warmed(False)
    ids, st = burst(srv)
    print(f"bench: served FP32 capacity {st['frames_per_s']:.0f} frames/s", file=sys.stderr, flush=True)
    if capacity_only:  # the warmed capacity pass alone (for kernel traces)
        return {"frames_per_s": st["frames_per_s"], "requests": st["completed"], "relayouts": st["relayouts"],
                "new_image_shells": st["reshells"], "max_batch_slots": st["max_slots"]}
    lat = poisson(srv)
    print(f"bench: served FP32 Poisson p50/p99 {lat['p50_ms']:.2f}/{lat['p99_ms']:.2f} ms", file=sys.stderr,
          flush=True)
    srv_bf = warmed(True)
    ids_bf, st_bf = burst(srv_bf)
    lat_bf = poisson(srv_bf)
    tunes = [n.tuning() for n in neps]
    # the plain path: one request at a time, FP32; single points go through
    # NEP.compute on device arrays, NVE segments through MD.run
    plain = [dmgx.NEP(p) for p in paths]
    errs = {"fp32": 0.0, "bf16": 0.0}

    def plain_one(k, check=False):
        r, model, steps = stream[k]
        n = len(r["types"])
        if steps == 0:
            t = torch.tensor(r["types"], dtype=torch.int32, device="cuda")
            x = torch.tensor(r["pos"].reshape(-1), dtype=torch.float64, device="cuda")
            pe = torch.zeros(n, dtype=torch.float64, device="cuda")
            f = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
            vir = torch.zeros(9 * n, dtype=torch.float64, device="cuda")
            plain[model].compute(r["box"], t, x, pe, f, vir)
            torch.cuda.synchronize()
            if check:
                fh = f.cpu().numpy().reshape(3, n)
                errs["fp32"] = max(errs["fp32"], float(np.abs(srv.poll(ids[k])["force"] - fh).max()))
                errs["bf16"] = max(errs["bf16"], float(np.abs(srv_bf.poll(ids_bf[k])["force"] - fh).max()))
            return 1
        md = dmgx.MD(plain[model], r["box"], r["types"], r["mass"], r["pos"], r["vel"])
        md.run(steps, r["dt_fs"], "nve")
        return steps

    for k in range(4):  # first launches of both models, single point and MD
        plain_one(k)
    frames = 0
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for k in range(requests):
        frames += plain_one(k)
    plain_fps = frames / (time.perf_counter() - tp)
    print(f"bench: plain FP32 {plain_fps:.0f} frames/s", file=sys.stderr, flush=True)
    for k in range(0, min(requests, 512), 2):  # single points: served against plain
        plain_one(k, check=True)
    n_carbon = sum(1 for _, m, _ in stream if m == 0)
    return {"workload": f"{requests} requests ({n_carbon} carbon diamond, {requests - n_carbon} Cu fcc, each with "
                        "its own NEP4 model), 64..20000 sites log-uniform, 30% vacancies on half of them, half "
                        "single points (energy, forces, virial) and half 50-step NVE (carbon 1 fs, Cu 2 fs)",
            "served_ann": "fp32 (matrix cores): the plain path's arithmetic, so equal force error",
            "frames_per_s": st["frames_per_s"], "atom_frames_per_s": st["atom_frames"] / max(st["elapsed_s"], 1e-9),
            "capacity_run": "all requests submitted at once by 8 client threads, on a server warmed by one "
                            "untimed pass of the stream",
            "requests": st["completed"], "relayouts": st["relayouts"], "iterations": st["iterations"],
            "new_image_shells": st["reshells"],
            "max_batch_slots": st["max_slots"],
            "hbm_budget_bytes": st["hbm_budget"], "max_bytes_in_flight": st["max_bytes_in_flight"],
            "worker_s": {"relayout": round(st["relayout_s"], 4), "iterate": round(st["iterate_s"], 4),
                         "retire": round(st["retire_s"], 4)},
            "latency_run": f"Poisson arrivals {rate_hz:.0f}/s", "p50_ms": lat["p50_ms"], "p99_ms": lat["p99_ms"],
            "latency_run_frames_per_s": lat["frames_per_s"],
            "plain_fp32_frames_per_s": plain_fps, "plain_requests": requests,
            "speedup_vs_plain": st["frames_per_s"] / plain_fps if plain_fps > 0 else None,
            "target_speedup": 8.0,
            "served_vs_plain_force_err_max_eV_per_A": errs["fp32"],
            "force_err_note": "served single points against the plain FP32 path on the stream's first 256 single "
                              "points; the plain path and the served FP32 mode are each held to the FP64 oracle "
                              "within 1e-4 eV/A by the GPU tests",
            "bf16": {"frames_per_s": st_bf["frames_per_s"], "p50_ms": lat_bf["p50_ms"], "p99_ms": lat_bf["p99_ms"],
                     "speedup_vs_plain": st_bf["frames_per_s"] / plain_fps if plain_fps > 0 else None,
                     "force_err_vs_plain_fp32_max_eV_per_A": errs["bf16"]},
            "tile_shapes": [{"model": os.path.basename(p), "filter_lanes_per_centre": t["filter_lanes"],
                             "gather_lanes_per_atom": t["gather_lanes"], "autotuned": t["tuned"],
                             "row_capacity": t["row_capacity"], "row_layout": t["row_layout"],
                             "rows_split_at_rebuild": t["split"]}
                            for p, t in zip(paths, tunes)]}


def counter_flops(kernel: str, atoms: int):
    """FP32 VALU and MFMA FLOP per atom of a kernel from the committed
    rocprofv3 counter summary, when it covers this workload; None otherwise."""
    try:
        d = json.load(open(os.path.join(ROOT, "profiles", "k_descriptor_cu256k_counters.json")))
    except (OSError, ValueError):
        return None
    if d.get("atoms") != atoms:
        return None
    return d["kernels"].get(kernel)


def small_bench(dmgx, steps=100, warmup=20):
    """The 4,096-atom carbon diamond configuration (NEP4 carbon model): NVE
    at dt 1 fs, `steps` timed steps after `warmup`."""
    import torch
    W = importlib.import_module("dmg-md_amd.workloads")
    s = W.carbon_diamond_4096(seed=1)
    N = len(s["types"])
    md = dmgx.MD(dmgx.NEP(CARBON_MODEL), s["box"], s["types"], s["mass"], s["pos"], s["vel"])
    md.run(warmup, 1.0, "nve")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    md.run(steps, 1.0, "nve")
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # the same loop replaying one captured step (bit-identical results)
    md.set_step_graph(True)
    md.run(warmup, 1.0, "nve")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    md.run(steps, 1.0, "nve")
    torch.cuda.synchronize()
    el_g = time.perf_counter() - t0
    return {"workload": "carbon diamond 8^3 cells (4096 atoms, a=3.567 A, 0.05 A jitter), NEP4 n_max 4/4 basis 8/8 "
                        "l_max 4/2/1 rc 4.2/3.7 A, 30 neurons; NVE 300 K, dt 1 fs", "atoms": N, "steps": steps,
            "atom_steps_per_s": N * steps / el, "ms_per_step": el / steps * 1e3,
            "ms_per_step_replayed_graph": el_g / steps * 1e3}


CARBON_ZBL_MODEL = os.path.join(ROOT, "tests", "golden", "potentials", "nep_carbon_zbl.txt")


def share_bench(dmgx, steps=200):
    """One GPU's share of an amorphous-carbon melt: 2,048,000 owned atoms at
    2.0 g/cm^3 plus a static 5 A ghost shell, NEP4-ZBL carbon model, NVE at
    dt 0.5 fs for the owned atoms (forces through the domain call: owned
    atoms as force centres, owned + ghosts as dependency centres; half kicks
    and drift as FP64 device array updates), FP32 and BF16 ANN operands."""
    import numpy as np
    import torch
    W = importlib.import_module("dmg-md_amd.workloads")
    t_gen = time.perf_counter()
    s = W.amorphous_carbon_share(2048000, seed=1)
    print(f"bench: a-C share generated in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)
    N, n = len(s["types"]), s["owned"]
    dt = 0.5 / 10.18051
    t = torch.tensor(s["types"], dtype=torch.int32, device="cuda")
    m = torch.tensor(s["mass"][:n], dtype=torch.float64, device="cuda")
    out = {"workload": f"amorphous carbon melt share: {n} owned atoms (2.0 g/cm^3, 1.2 A minimum separation, "
                       f"5000 K) + {N - n} static ghost-shell atoms (5 A), NEP4-ZBL carbon model (ZBL 1.0/2.0 A), "
                       "NVE dt 0.5 fs of the owned atoms", "owned_atoms": n, "slots": N, "steps": steps}
    for mode in ("fp32", "bf16"):
        nep = dmgx.NEP(CARBON_ZBL_MODEL)
        x = torch.tensor(np.ascontiguousarray(s["pos"]).reshape(-1), dtype=torch.float64, device="cuda")
        xv = x.view(3, N)
        v = torch.tensor(np.ascontiguousarray(s["vel"]), dtype=torch.float64, device="cuda").view(3, n)
        pe = torch.zeros(N, dtype=torch.float64, device="cuda")
        f = torch.zeros(3 * N, dtype=torch.float64, device="cuda")
        vir = torch.zeros(9 * N, dtype=torch.float64, device="cuda")
        fv = f.view(3, N)

        def force():
            nep.compute_domain(s["box"], N, t, x, pe, f, vir, n1=0, n2=n, nd1=0, nd2=N, ann_bf16=mode == "bf16")

        def step():
            v.add_(fv[:, :n] / m, alpha=0.5 * dt)
            xv[:, :n].add_(v, alpha=dt)
            force()
            v.add_(fv[:, :n] / m, alpha=0.5 * dt)

        force()
        for _ in range(5):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[mode] = {"atom_steps_per_s": n * steps / el, "ms_per_step": el / steps * 1e3,
                     "verlet_rebuilds": nep.neighbor_stats()["rebuilds"]}
        del nep
    return out


def cpu_baseline_parallel(procs=8, budget_s=10.0):
    """The same CPU sample in `procs` independent processes at once (one per
    core): the host's aggregate FP64 throughput."""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    with ctx.Pool(procs) as pool:
        res = pool.map(_cpu_worker, [budget_s] * procs)
    return {"value": sum(r["value"] for r in res), "unit": "atom-steps/s", "cores": procs, "kind": res[0]["kind"],
            "sample": f"{procs} processes, each running the 1-core sample independently; aggregate throughput"}


def _cpu_worker(budget_s):
    return cpu_baseline(budget_s)


def host_cpu():
    """CPU model and the cores this process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu_model": model, "cores_available": avail, "nproc": os.cpu_count()}


def context_bench(dmgx):
    """The reference's only published timing job, replayed through the
    run.in front end (dmgx_run_directory): 5,760-atom diamond carbon (20x6x6
    cells), the reference's nep_C.txt (NEP4, rc 7/4 A, n_max 12/8, basis
    16/12, 100 neurons), npt_ber 300 K / 0 GPa (Tdamp 40 steps, B 425 GPa,
    tau_p 1000 steps), dt 0.1 fs, thermo every 100 steps, a restart at the
    end.  The steady-state step time is the difference of a 2,000-step and
    a 1,000-step job's wall times over 1,000 steps (the published figure is
    the mean over steps 1,000-2,000 of a 2,000-step job)."""
    import shutil
    import tempfile
    W = importlib.import_module("dmg-md_amd.workloads")
    cli = importlib.import_module("dmg-md_amd.cli")
    model = W.diamond_npt_model_xyz((20, 6, 6), 301)
    pot = os.path.join(ROOT, "tests", "golden", "baseline", "inputs", "potentials", "nep_C.txt")
    walls = {}
    for steps in (1000, 2000):
        d = tempfile.mkdtemp(prefix="dmgx_ctx_")
        try:
            open(os.path.join(d, "model.xyz"), "w").write(model)
            shutil.copy(pot, os.path.join(d, "nep.txt"))
            open(os.path.join(d, "run.in"), "w").write(
                "potential nep.txt\ntime_step 0.1\nensemble npt_ber 300 300 40 0 425 1000\n"
                f"dump_thermo 100\ndump_restart {steps}\nrun {steps}\n")
            t0 = time.perf_counter()
            cli.run_directory(d)
            walls[steps] = time.perf_counter() - t0
            rows = open(os.path.join(d, "thermo.out")).read().split("\n")
            last = [float(x) for x in rows[-2].split()] if len(rows) > 1 else []
        finally:
            shutil.rmtree(d, ignore_errors=True)
    ms = (walls[2000] - walls[1000]) / 1000 * 1e3
    return {"workload": "5,760-atom diamond carbon (20x6x6 cells, a = 3.57 x 1.0023 A), the reference's nep_C.txt, "
                        "npt_ber 300 K / 0 GPa, Tdamp 40, B 425 GPa, tau_p 1000, dt 0.1 fs, thermo every 100 steps, "
                        "restart at the end, through dmgx_run_directory",
            "wall_s_1000_steps": walls[1000], "wall_s_2000_steps": walls[2000],
            "steady_ms_per_step": ms, "atom_steps_per_s": 5760 / (ms * 1e-3),
            "final_thermo_row": last,
            "published_L20_ms_per_step": {"DMG-MD (1 rank, HostStaged)": 5.1768, "GPUMD 5.9.1": 3.5262},
            "published_source": "docs/status/performance/npt-benchmarks.md:52-57 (NVIDIA L20)"}


WATER_MODEL = os.path.join(ROOT, "tests", "golden", "potentials", "nep_water.txt")


def water_bench(dmgx, steps=100):
    """The liquid-water configuration (64,000 H2O, 192,000 atoms, 2-type
    NEP4 n_max 6/6, basis 10, l_max 4/2/1, rc 6/4 A, 40 neurons), dt 0.5 fs
    at 300 K: a Bussi NVT stretch (the sync-free chunked step) and the
    configuration's SCR NPT (1e-4 GPa; the barostat reads the pressure on
    the host every step), each after 20 untimed steps. The generator inserts
    molecules at random, so their storage order is spatially random: both
    runs store the atoms in cell order first (dmgx_md_reorder), and the NVT
    run is also timed in the generator's order."""
    import torch
    W = importlib.import_module("dmg-md_amd.workloads")
    s = W.liquid_water(64000, seed=1)
    N = len(s["types"])
    out = {"workload": "liquid water 64,000 H2O (192,000 atoms), 2-type NEP4 n_max 6/6 basis 10 l_max 4/2/1 "
                       "rc 6.0/4.0 A 40 neurons, dt 0.5 fs, 300 K", "atoms": N,
           "storage_order": "cell order (dmgx_md_reorder) unless the key says input order"}
    for name, ens, kw, reorder in (
            ("nvt_bdp", "nvt_bdp", dict(t_coupling=100.0), True),
            ("npt_scr", "npt_scr", dict(t_coupling=100.0, p_target_gpa=1e-4, bulk_modulus_gpa=2.2, p_coupling=1000.0),
             True),
            ("nvt_bdp_input_order", "nvt_bdp", dict(t_coupling=100.0), False)):
        nep = dmgx.NEP(WATER_MODEL)
        md = dmgx.MD(nep, s["box"], s["types"], s["mass"], s["pos"], s["vel"])
        if reorder:
            md.reorder()
        tune = md.autotune()
        md.run(20, 0.5, ens, t_init=300.0, t_final=300.0, **kw)
        torch.cuda.synchronize()
        r0 = nep.neighbor_stats()["rebuilds"]
        t0 = time.perf_counter()
        md.run(steps, 0.5, ens, t_init=300.0, t_final=300.0, **kw)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[name] = {"atom_steps_per_s": N * steps / el, "ms_per_step": el / steps * 1e3, "steps": steps,
                     "verlet_rebuilds": nep.neighbor_stats()["rebuilds"] - r0,
                     "tile_shapes": {k: tune[k] for k in ("filter_lanes", "gather_lanes", "row_layout", "overlap",
                                                           "split", "desc_lanes")}}
        if ens == "npt_scr":
            # averages over a further stretch, read every 5 steps
            vols, press = [], []
            for _ in range(20):
                th = md.run(5, 0.5, ens, t_init=300.0, t_final=300.0, **kw)
                vols.append(md.box().volume)
                press.append(float(th[2] + th[3] + th[4]) / 3.0 * dmgx.PRESSURE_UNIT_CONVERSION)
            out[name].update({"mean_volume_A3": float(sum(vols) / len(vols)),
                              "initial_volume_A3": s["box"].volume,
                              "mean_pressure_GPa": float(sum(press) / len(press)),
                              "averaged_over": "100 further steps, sampled every 5"})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cells", type=int, default=40, help="fcc cells per edge (40 -> 256,000 atoms)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop sample")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default filter/gather tile shapes")
    ap.add_argument("--no-water", action="store_true", help="skip the liquid-water (2-type) sample")
    ap.add_argument("--no-small", action="store_true", help="skip the 4,096-atom carbon NVE line")
    ap.add_argument("--no-context", action="store_true",
                    help="skip the replay of the reference's published 5,760-atom NPT timing job")
    ap.add_argument("--no-share", action="store_true", help="skip the 2,048,000-atom amorphous-carbon share line")
    ap.add_argument("--amortised-steps", type=int, default=300,
                    help="steps of the stretch timed with its Verlet rebuilds")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving sample (its JSON object)")
    ap.add_argument("--serve-capacity-only", action="store_true",
                    help="with --serve-only: only the warm-up and the timed FP32 capacity pass")
    ap.add_argument("--water-only", type=int, default=0, metavar="STEPS",
                    help="run only the liquid-water sample with STEPS timed steps per run and print its "
                         "JSON object (for kernel profiles of the 2-type path)")
    ap.add_argument("--rows", choices=["auto", "filter", "split"], default="auto",
                    help="typewise row split every step (filter), once per rebuild (split), or the autotuner's")
    ap.add_argument("--ann", choices=["fp32", "valu", "bf16"], default="fp32",
                    help="ANN arithmetic: fp32 on the matrix cores (default), fp32 on the vector ALUs, "
                         "or bfloat16 operands on the matrix cores")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB five times (the known-size "
                         "kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 * n, "launches": 5}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(f"bench.py measures one GPU's share of the domain; --gpus {max(args.gpus, world)} is not "
              "supported (multi-GPU halo exchange is out of scope for this build)", file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        sys.exit("bench.py needs a GPU")
    dmgx = importlib.import_module("dmg-md_amd")
    if args.water_only > 0:
        print(json.dumps(water_bench(dmgx, steps=args.water_only)))
        return
    if args.serve_only:
        print(json.dumps(serve_bench(dmgx, capacity_only=args.serve_capacity_only)))
        return
    W = importlib.import_module("dmg-md_amd.workloads")
    s = W.copper_fcc(n=args.cells, seed=1)
    N = len(s["types"])
    nep = dmgx.NEP(MODEL)
    md = dmgx.MD(nep, s["box"], s["types"], s["mass"], s["pos"], s["vel"])
    ens = dict(ensemble="nvt_bdp", t_init=300.0, t_final=300.0, t_coupling=100.0, seed=12345678,
               ann_bf16=args.ann == "bf16", ann_valu=args.ann == "valu")
    print(f"bench: Cu fcc {args.cells}^3 = {N} atoms, warmup {args.warmup} steps", file=sys.stderr, flush=True)
    # filter / gather tile shapes timed on this system before anything else
    tune = nep.tuning() if args.no_autotune else md.autotune(ann_bf16=args.ann == "bf16")
    if args.rows != "auto":
        nep.set_split(args.rows == "split")
        tune = nep.tuning()
    print(f"bench: tile shapes {tune}", file=sys.stderr, flush=True)
    md.run(args.warmup, 2.0, **ens)
    torch.cuda.synchronize()
    # timed region: only the dominant kernel is timed, with HIP events
    # attached to its own dispatch (no extra packets in the stream)
    nep.profile_reset()
    nep.set_profiling(True, dispatch_only=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    th = md.run(args.steps, 2.0, **ens)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    nep.set_profiling(False)
    prof = nep.profile()
    means = nep.neighbor_means()
    stats = nep.neighbor_stats()
    elapsed = t1 - t0
    value = N * args.steps / elapsed
    fl = flops_per_atom(nep.info, means["radial"], means["angular"])
    n_fused, ms_fused = prof["descriptor"]
    avg_fused = ms_fused / max(n_fused, 1)
    # the fused kernel's time per step (one launch over all atoms, or two
    # over half of them each with the overlapped filter schedule)
    per_step_fused = ms_fused / args.steps
    achieved = fl["fused"] * N / (per_step_fused * 1e-3) / 1e12
    # per-kernel breakdown from a separate, untimed profiled stretch (events
    # around every kernel group separate the kernels by a few microseconds)
    nep.profile_reset()
    nep.set_profiling(True)
    md.run(min(args.steps, 100), 2.0, **ens)
    torch.cuda.synchronize()
    nep.set_profiling(False)
    kernel_ms = {k: round(v[1] / max(v[0], 1), 4) for k, v in nep.profile().items()}
    print(f"bench: {args.steps} steps in {elapsed:.3f} s, T={th[0]:.1f} K, rebuilds={stats['rebuilds']}, "
          f"per-kernel avg ms {kernel_ms}, mean rows {means}", file=sys.stderr, flush=True)
    # a longer stretch that holds Verlet rebuilds at their natural rate
    # (the timed steps above may hold none when K is small)
    r0 = nep.neighbor_stats()["rebuilds"]
    torch.cuda.synchronize()
    ta = time.perf_counter()
    md.run(args.amortised_steps, 2.0, **ens)
    torch.cuda.synchronize()
    el_a = time.perf_counter() - ta
    rebuilds_a = nep.neighbor_stats()["rebuilds"] - r0
    amortised = {"steps": args.amortised_steps, "verlet_rebuilds": rebuilds_a,
                 "ms_per_step": el_a / args.amortised_steps * 1e3,
                 "atom_steps_per_s": N * args.amortised_steps / el_a}
    # the same configuration with bfloat16 ANN operands, fresh state
    # the same configuration with bfloat16 ANN operands, and with split
    # bfloat16 operands (three parts each, FP32 accuracy), fresh states
    bf16 = bf16_split = None
    if args.ann == "fp32":
        def ann_line(kw, label):
            nep_b = dmgx.NEP(MODEL)
            md_b = dmgx.MD(nep_b, s["box"], s["types"], s["mass"], s["pos"], s["vel"])
            md_b.autotune(**kw)
            ens_b = dict(ens, **kw)
            md_b.run(args.warmup, 2.0, **ens_b)
            torch.cuda.synchronize()
            tb = time.perf_counter()
            md_b.run(args.steps, 2.0, **ens_b)
            torch.cuda.synchronize()
            el_b = time.perf_counter() - tb
            return {"ann": label, "steps": args.steps, "atom_steps_per_s": N * args.steps / el_b,
                    "ms_per_step": el_b / args.steps * 1e3}
        bf16 = ann_line({"ann_bf16": True}, "bf16 operands, fp32 accumulation (matrix cores)")
        bf16_split = ann_line({"ann_bf16_split": True},
                              "split bf16 operands (hi + mid + lo, six products), fp32 accumulation (matrix cores)")
    small = None if args.no_small else small_bench(dmgx)
    share = None if args.no_share else share_bench(dmgx)
    context = None if args.no_context else context_bench(dmgx)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_seconds)
        cpu.update(host_cpu())
        cpu["parallel"] = cpu_baseline_parallel(8, args.cpu_seconds)
    serving = None if args.no_serve else serve_bench(dmgx)
    water = None if args.no_water else water_bench(dmgx)
    # the fused kernel's measured arithmetic (committed counter summary of
    # this workload) over its live launch time
    cnt = counter_flops("k_descriptor", N)
    valu_measured = None
    if cnt:
        valu_measured = {"valu_fp32_flop_per_atom": cnt["valu_fp32_flop_per_atom"],
                         "mfma_flop_per_atom": cnt["mfma_flop_per_atom"],
                         "valu_tflops": cnt["valu_fp32_flop_per_atom"] * N / (per_step_fused * 1e-3) / 1e12,
                         "scalar_fp32_peak_tflops": PEAK_FP32_SCALAR_TFLOPS,
                         "frac_of_scalar_peak": cnt["valu_fp32_flop_per_atom"] * N / (per_step_fused * 1e-3) / 1e12
                         / PEAK_FP32_SCALAR_TFLOPS,
                         "valu_issue_share": cnt["valu_issue_share"],
                         "source": "profiles/k_descriptor_cu256k_counters.md"}
    # the memory-bound per-step kernels against HBM: algorithmic bytes per
    # atom (the filter's from the measurement plan's formula, the gather's
    # minimal streams) and the counters' bytes, over the profiled stretch's
    # average launch time
    n_r, n_a, n_v = means["radial"], means["angular"], means["verlet"]
    kRp = (nep.info.basis_size_radial + 1 + 3) & ~3
    alg = {"k_filter": 4 * n_v + 4 * (n_r + n_a) + 8,
           "k_gather": 16 * n_r + 24 * n_a + 4 * kRp * nep.info.num_types + 96}
    hbm = {}
    for k, key in (("k_filter", "filter"), ("k_gather", "gather")):
        t_ms = kernel_ms.get(key, 0.0)
        if t_ms <= 0:
            continue
        tr = measured_traffic("dmgx::" + k, N)
        e = {"algorithmic_bytes_per_atom": round(alg[k], 1), "avg_launch_ms": t_ms,
             "algorithmic_GBs": alg[k] * N / (t_ms * 1e-3) / 1e9}
        e["frac_algorithmic"] = e["algorithmic_GBs"] / PEAK_HBM_GBS
        if tr:
            e["counter_bytes_per_atom"] = round(tr / N, 1)
            e["counter_GBs"] = tr / (t_ms * 1e-3) / 1e9
            e["frac_counter"] = e["counter_GBs"] / PEAK_HBM_GBS
        hbm[k] = e
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 force/descriptor/ANN, f64 positions/integrator/thermo" if args.ann != "bf16" else
                  "f32 force/descriptor, bf16 ANN operands, f64 positions/integrator/thermo"),
        "data": "synthetic (seeded Cu fcc lattice + 0.03 A jitter, seeded random NEP4 weights)",
        "config": {
            "workload": f"Cu fcc {args.cells}^3 cells ({N} atoms, a=3.615 A), NEP4 n_max 6/4 basis 8/8 "
                        "l_max 4/2/0 rc 5.0/4.0 A, 50 tanh neurons; Bussi NVT 300 K, tau 100 steps, dt 2 fs",
            "atoms": N, "ensemble": "nvt_bdp", "dt_fs": 2.0,
            "ann": {"fp32": "fp32 (matrix cores)", "valu": "fp32 (vector ALUs)",
                    "bf16": "bf16 operands, fp32 accumulation (matrix cores)"}[args.ann],
            "mean_neighbors": {k: round(v, 2) for k, v in means.items() if k != "centres"},
            "verlet_rebuilds_in_run": stats["rebuilds"],
            "kernel_avg_ms": kernel_ms,
            "kernel_avg_ms_from": "a separate profiled stretch after the timed steps",
            "tile_shapes": {"filter_lanes_per_centre": tune["filter_lanes"],
                            "gather_lanes_per_atom": tune["gather_lanes"],
                            "autotuned": tune["tuned"],
                            "filter_ms_by_lanes": {str(k): round(v, 4) for k, v in tune["filter_ms"].items()},
                            "gather_ms_by_lanes": {str(k): round(v, 4) for k, v in tune["gather_ms"].items()},
                            "row_capacity": tune["row_capacity"],
                            "row_layout": {0: "multiple of 4", 1: "power of 2"}.get(tune["row_layout"], "default"),
                            "force_ms_by_row_layout": {k: round(v, 4) for k, v in tune["force_ms"].items()},
                            "filter_descriptor_overlap": tune["overlap"],
                            "force_ms_by_schedule": {k: round(v, 4) for k, v in tune["overlap_ms"].items()},
                            "rows_split_at_rebuild": tune["split"],
                            "force_ms_by_row_split": {k: round(v, 4) for k, v in tune["split_ms"].items()}},
            "final_temperature_K": th[0],
        },
        "ms_per_step_with_rebuilds": amortised,
        "bf16_ann": bf16,
        "bf16_split_ann": bf16_split,
        "roofline": {
            "kernel": "k_descriptor (fused descriptor + ANN + adjoint + angular partial forces)",
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": measured_traffic("dmgx::k_descriptor", N),
            "flop_per_atom": fl["fused"],
            "flop_count": "algorithmic: the reference algorithm's FLOP per atom (measurement-plan formulas)",
            "avg_launch_ms": avg_fused,
            "launches_per_step": n_fused / args.steps,
            "ms_per_step_in_kernel": per_step_fused,
            "measured_valu": valu_measured,
            "hbm_kernels": hbm,
        },
        "cpu_baseline": cpu,
        "serving": serving,
        "carbon_4096": small,
        "carbon_share": share,
        "published_job": context,
        "water": water,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
