"""The MD loop replaying one captured step (dmgx_md_set_step_graph: a
hipGraph of the fused first half, the force kernels, the second half and the
thermostat reduction) instead of launching each kernel.  Through the C ABI:
the replayed runs equal the directly launched ones bit for bit, over chunk
boundaries, Verlet rebuilds, a Bussi thermostat's per-step draws and a
Berendsen temperature ramp (per-step targets), and over replays with the
host's own work between them; the step times of both are printed."""
import time

import numpy as np
import pytest

from test_gpu_parity import CARBON

pytestmark = pytest.mark.gpu


def _run(dmgx, s, vel, steps, direct, **ens):
    nep = dmgx.NEP(CARBON)
    md = dmgx.MD(nep, s["box"], s["types"], s["mass"], s["pos"], vel)
    md.set_step_graph(not direct)
    th = []
    t0 = time.perf_counter()
    for n in steps:  # several runs in a row: host work between the replays
        th.append(md.run(n, 1.0, **ens))
    ms = (time.perf_counter() - t0) / sum(steps) * 1e3
    print(f"{'direct launches' if direct else 'replayed step'}: {ms:.4f} ms/step (first runs included)", flush=True)
    return md.arrays(), np.array(th), nep.neighbor_stats()["rebuilds"]


@pytest.mark.parametrize("ens", [
    dict(ensemble="nve"),
    dict(ensemble="nvt_bdp", t_init=3000.0, t_final=3000.0, t_coupling=20.0, seed=7),
    dict(ensemble="nvt_ber", t_init=3000.0, t_final=1000.0, t_coupling=20.0),
], ids=["nve", "nvt_bdp", "nvt_ber_ramp"])
def test_replayed_step_equals_direct_launches(dmgx, workloads, ens):
    s = workloads.carbon_diamond_4096(seed=21)
    # hot (3,000 K): the rows are rebuilt several times; 150 + 40 steps
    # cross a chunk boundary (128 steps) and start a second run
    vel = np.ascontiguousarray(s["vel"]) * np.sqrt(10.0)
    a, tha, ra = _run(dmgx, s, vel, (150, 40), direct=True, **ens)
    b, thb, rb = _run(dmgx, s, vel, (150, 40), direct=False, **ens)
    print(f"{ens['ensemble']}: rebuilds {ra} / {rb}, T {tha[-1][0]:.3f} / {thb[-1][0]:.3f} K", flush=True)
    assert ra == rb and ra >= 3
    for k in ("position", "velocity", "force", "potential", "virial"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(tha, thb)
