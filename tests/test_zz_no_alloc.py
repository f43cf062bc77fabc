"""dmgx_nep_compute_noalloc, the evaluation whose workspace is the library's
fixed device array, against dmgx_nep_compute (the allocating entry point) on
the same handle and inputs, both called through the C ABI: two calls in a
row, the second on a smaller system with domain ranges, the fixed array
keeping the first call's contents in between.  The two entry points run the
same kernels on the same rows, so the outputs agree to rounding (bitwise in
practice); the bar written here is 1e-9 eV/A, eV and eV per virial
component, far inside the 1e-4 eV/A force parity bar."""
import ctypes
import importlib
import os

import numpy as np
import pytest

from conftest import POTENTIALS

pytestmark = pytest.mark.gpu

CARBON = os.path.join(POTENTIALS, "nep_carbon.txt")
TOL = 1e-9


def _call(dmgx, entry, nep, s, ranges=None, flags=0, gid=None):
    import torch
    L = importlib.import_module("dmg-md_amd._lib")
    N = len(s["types"])
    n1, n2, nd1, nd2 = ranges or (0, N, 0, N)
    t = torch.tensor(s["types"], dtype=torch.int32, device="cuda")
    x = torch.tensor(np.ascontiguousarray(s["pos"]).reshape(-1), dtype=torch.float64, device="cuda")
    g = None if gid is None else torch.tensor(gid, dtype=torch.int64, device="cuda")
    # outputs start from NaN: every force centre's values must be written
    pe = torch.full((N,), float("nan"), dtype=torch.float64, device="cuda")
    f = torch.full((3 * N,), float("nan"), dtype=torch.float64, device="cuda")
    v = torch.full((9 * N,), float("nan"), dtype=torch.float64, device="cuda")
    b = s["box"].to_c()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    fn = getattr(L.lib(), entry)
    status = fn(nep.handle, ctypes.byref(b), N, n1, n2, nd1, nd2, ctypes.c_void_p(t.data_ptr()),
                ctypes.c_void_p(x.data_ptr()), None if g is None else ctypes.c_void_p(g.data_ptr()),
                ctypes.c_void_p(pe.data_ptr()), ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(v.data_ptr()),
                flags, stream)
    L.check(status)
    torch.cuda.synchronize()
    return (pe.cpu().numpy()[n1:n2], f.cpu().numpy().reshape(3, N)[:, n1:n2],
            v.cpu().numpy().reshape(9, N)[:, n1:n2])


def _agree(a, b, what):
    for name, x, y in zip(("potential", "force", "virial"), a, b):
        assert np.isfinite(x).all() and np.isfinite(y).all(), (what, name)
        d = float(np.abs(x - y).max())
        print(f"{what}: {name} max |noalloc - compute| = {d:.3e}", flush=True)
        assert d <= TOL, (what, name, d)


def test_noalloc_matches_compute_on_two_calls(dmgx, workloads):
    nep = dmgx.NEP(CARBON)
    big = workloads.carbon_diamond_4096(seed=5)
    print("no-alloc: carbon diamond 4096 atoms, all centres, fp32 ANN", flush=True)
    a = _call(dmgx, "dmgx_nep_compute_noalloc", nep, big)
    b = _call(dmgx, "dmgx_nep_compute", nep, big)
    _agree(a, b, "4096 atoms")
    # a smaller system next (the fixed array still holds the first call's
    # rows and cell counts), with force centres inside the dependency centres
    # and rows keyed by global id
    pos, box = workloads.diamond(6, 3.567)
    pos = workloads.wrap(workloads.jitter(pos, 0.05, 9), box)
    small = dict(box=box, pos=workloads.soa(pos), types=np.zeros(pos.shape[0], np.int32))
    N = len(small["types"])
    gid = np.random.default_rng(3).permutation(N).astype(np.int64)
    ranges = (200, 1400, 100, 1600)
    print(f"no-alloc: carbon diamond {N} atoms, centres {ranges}, rows by global id", flush=True)
    a = _call(dmgx, "dmgx_nep_compute_noalloc", nep, small, ranges, flags=8, gid=gid)
    b = _call(dmgx, "dmgx_nep_compute", nep, small, ranges, flags=8, gid=gid)
    _agree(a, b, f"{N} atoms, domain ranges")
    # the handle's own list and workspace were left alone by the fixed-array
    # calls: the allocating path still gives the same first result
    c = _call(dmgx, "dmgx_nep_compute", nep, big)
    a = _call(dmgx, "dmgx_nep_compute_noalloc", nep, big)
    _agree(a, c, "4096 atoms again")


def test_noalloc_rejects_what_does_not_fit(dmgx, workloads):
    """A system beyond the fixed array is refused before any kernel runs,
    with the sizes in the message; so is a replicated small box."""
    L = importlib.import_module("dmg-md_amd._lib")
    nep = dmgx.NEP(CARBON)
    pos, box = workloads.diamond(20, 3.567)  # 64,000 atoms
    s = dict(box=box, pos=workloads.soa(pos), types=np.zeros(pos.shape[0], np.int32))
    with pytest.raises(L.DmgxError) as e:
        _call(dmgx, "dmgx_nep_compute_noalloc", nep, s)
    assert e.value.code == -1 and "fixed workspace holds" in str(e.value)
    pos, box = workloads.diamond(1, 3.567)  # 8 atoms in a 3.567 A cube
    s = dict(box=box, pos=workloads.soa(pos), types=np.zeros(pos.shape[0], np.int32))
    with pytest.raises(L.DmgxError) as e:
        _call(dmgx, "dmgx_nep_compute_noalloc", nep, s)
    assert e.value.code == -1 and "replicated" in str(e.value)


def test_noalloc_with_candidate_rows(dmgx, workloads):
    """The same two calls on a handle whose rows are split at rebuild time
    (the fixed array then also holds the candidate rows)."""
    nep = dmgx.NEP(CARBON)
    nep.set_split(True)
    big = workloads.carbon_diamond_4096(seed=6)
    print("no-alloc, candidate rows: carbon diamond 4096 atoms", flush=True)
    a = _call(dmgx, "dmgx_nep_compute_noalloc", nep, big)
    b = _call(dmgx, "dmgx_nep_compute", nep, big)
    _agree(a, b, "4096 atoms, candidate rows")
    pos, box = workloads.diamond(6, 3.567)
    pos = workloads.wrap(workloads.jitter(pos, 0.05, 10), box)
    small = dict(box=box, pos=workloads.soa(pos), types=np.zeros(pos.shape[0], np.int32))
    ranges = (100, 1500, 0, 1728)
    print(f"no-alloc, candidate rows: carbon diamond 1728 atoms, centres {ranges}", flush=True)
    a = _call(dmgx, "dmgx_nep_compute_noalloc", nep, small, ranges)
    b = _call(dmgx, "dmgx_nep_compute", nep, small, ranges)
    _agree(a, b, "1728 atoms, candidate rows")
