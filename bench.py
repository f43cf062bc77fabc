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


def _carbon_request(rng, n_min=64, n_max=20000):
    """A carbon diamond supercell of log-uniform size in [n_min, n_max]
    atoms (cubic n^3 cells), 30% vacancies, 0.05 A jitter."""
    a = 3.567
    frac = [[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
            [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]]
    import numpy as np
    target = float(np.exp(rng.uniform(np.log(n_min / 0.7), np.log(n_max / 0.7))))
    n = max(2, int(round((target / 8) ** (1 / 3))))
    cells = np.array([[i, j, k] for i in range(n) for j in range(n) for k in range(n)])
    pos = ((cells[:, None, :] + np.array(frac)[None]).reshape(-1, 3) * a)
    keep = rng.random(len(pos)) >= 0.3
    pos = pos[keep].T + rng.normal(0, 0.05, (3, int(keep.sum())))
    N = pos.shape[1]
    vel = rng.normal(0, 0.01, (3, N))
    vel -= vel.mean(1, keepdims=True)
    return np.diag([a * n] * 3), pos, vel


def serve_bench(dmgx, requests=256, rate_hz=2000.0, plain_requests=24, seed=7):
    """Served frames/s and latency percentiles of the continuous-batching
    server on a bounded sample of the serving workload (carbon requests of
    64..20000 atoms with 30% vacancies, half single points and half 50-step
    NVE, BF16 ANN).  Capacity: the whole stream submitted at once (closed
    loop, like the plain loop it is compared with).  Latency: the same
    stream again with Poisson arrivals at rate_hz.  Beside them a plain
    sequential FP32 loop (one request at a time through NEP.compute /
    MD.run) on the first requests of the stream, and the BF16 force error
    against FP32."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(requests):
        h, p, v = _carbon_request(rng)
        reqs.append((h, p, v, 50 if k % 2 else 0))
    gaps = rng.exponential(1.0 / rate_hz, requests)

    def submit(srv, k):
        h, p, v, stp = reqs[k]
        n = p.shape[1]
        return srv.submit(dmgx.Box(h), np.zeros(n, np.int32), np.full(n, 12.011), p, v, steps=stp, dt_fs=0.5)

    def burst(count, nep, clients=8):
        # `clients` threads submit the stream at once (the submit call
        # prepares a request's image shell and upload on the calling thread)
        from concurrent.futures import ThreadPoolExecutor
        srv = dmgx.Server(nep, max_atoms=6_000_000, ann_bf16=True)
        with ThreadPoolExecutor(clients) as pool:
            ids = list(pool.map(lambda k: submit(srv, k), range(count)))
        srv.drain()
        return srv, ids, srv.stats()

    def poisson(nep):
        srv = dmgx.Server(nep, max_atoms=6_000_000, ann_bf16=True)
        t_next = time.perf_counter()
        for k, g in enumerate(gaps):
            t_next += g
            while time.perf_counter() < t_next:
                time.sleep(min(1e-4, max(0.0, t_next - time.perf_counter())))
            submit(srv, k)
        srv.drain()
        return srv.stats()

    # first launches of the serving kernels; the server's autotuner picks the
    # tile shapes on the first batch of >= 64k atoms, here in this warm-up
    nep_srv = dmgx.NEP(CARBON_MODEL)
    burst(min(requests, 48), nep_srv)
    tune = nep_srv.tuning()
    srv, ids, st = burst(requests, nep_srv)
    lat = poisson(nep_srv)
    # BF16 force error against FP32 on single points of the stream
    err = 0.0
    nep32 = dmgx.NEP(CARBON_MODEL)
    for k in range(0, min(requests, 16), 2):
        h, p, v, _ = reqs[k]
        n = p.shape[1]
        got = srv.poll(ids[k])
        t = torch.zeros(n, dtype=torch.int32, device="cuda")
        x = torch.tensor(p.reshape(-1), dtype=torch.float64, device="cuda")
        pe = torch.zeros(n, dtype=torch.float64, device="cuda")
        f = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
        vir = torch.zeros(9 * n, dtype=torch.float64, device="cuda")
        nep32.compute(dmgx.Box(h), t, x, pe, f, vir)
        torch.cuda.synchronize()
        err = max(err, float(np.abs(got["force"] - f.cpu().numpy().reshape(3, n)).max()))
    # the plain path: one request at a time, FP32, after one untimed request
    # (first launches of its kernels)
    def plain_one(h, p, v, stp):
        n = p.shape[1]
        if stp == 0:
            t = torch.zeros(n, dtype=torch.int32, device="cuda")
            x = torch.tensor(p.reshape(-1), dtype=torch.float64, device="cuda")
            pe = torch.zeros(n, dtype=torch.float64, device="cuda")
            f = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
            vir = torch.zeros(9 * n, dtype=torch.float64, device="cuda")
            nep32.compute(dmgx.Box(h), t, x, pe, f, vir)
            torch.cuda.synchronize()
            return 1
        md = dmgx.MD(nep32, dmgx.Box(h), np.zeros(n, np.int32), np.full(n, 12.011), p, v)
        md.run(stp, 0.5, "nve")
        return stp

    plain_one(*reqs[1])
    frames = 0
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for h, p, v, stp in reqs[:plain_requests]:
        frames += plain_one(h, p, v, stp)
    plain = frames / (time.perf_counter() - tp)
    return {"workload": f"{requests} carbon requests, 64..20000 atoms (log-uniform, 30% vacancies), "
                        "half single points / half 50-step NVE at 0.5 fs, carbon NEP4 model, bf16 ANN",
            "frames_per_s": st["frames_per_s"], "atom_frames_per_s": st["atom_frames"] / max(st["elapsed_s"], 1e-9),
            "capacity_run": "all requests submitted at once by 8 client threads",
            "requests": st["completed"], "relayouts": st["relayouts"], "iterations": st["iterations"],
            "max_batch_slots": st["max_slots"],
            "worker_s": {"relayout": round(st["relayout_s"], 4), "iterate": round(st["iterate_s"], 4),
                         "retire": round(st["retire_s"], 4)},
            "latency_run": f"Poisson arrivals {rate_hz:.0f}/s", "p50_ms": lat["p50_ms"], "p99_ms": lat["p99_ms"],
            "latency_run_frames_per_s": lat["frames_per_s"],
            "bf16_force_err_max_eV_per_A": err,
            "tile_shapes": {"filter_lanes_per_centre": tune["filter_lanes"],
                            "gather_lanes_per_atom": tune["gather_lanes"], "autotuned": tune["tuned"],
                            "row_capacity": tune["row_capacity"], "row_layout": tune["row_layout"]},
            "plain_fp32_frames_per_s": plain, "plain_sample_requests": plain_requests,
            "speedup_vs_plain": st["frames_per_s"] / plain if plain > 0 else None}


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
        md.autotune()
        md.run(20, 0.5, ens, t_init=300.0, t_final=300.0, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        md.run(steps, 0.5, ens, t_init=300.0, t_final=300.0, **kw)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[name] = {"atom_steps_per_s": N * steps / el, "ms_per_step": el / steps * 1e3, "steps": steps}
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
    ap.add_argument("--water-only", type=int, default=0, metavar="STEPS",
                    help="run only the liquid-water sample with STEPS timed steps per run and print its "
                         "JSON object (for kernel profiles of the 2-type path)")
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
    achieved = fl["fused"] * N / (avg_fused * 1e-3) / 1e12
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
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_seconds)
    serving = None if args.no_serve else serve_bench(dmgx)
    water = None if args.no_water else water_bench(dmgx)
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
                            "force_ms_by_row_layout": {k: round(v, 4) for k, v in tune["force_ms"].items()}},
            "final_temperature_K": th[0],
        },
        "roofline": {
            "kernel": "k_descriptor (fused descriptor + ANN + adjoint + angular partial forces)",
            "bound": "mfma",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": measured_traffic("dmgx::k_descriptor", N),
            "flop_per_atom": fl["fused"],
            "avg_launch_ms": avg_fused,
        },
        "cpu_baseline": cpu,
        "serving": serving,
        "water": water,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
