"""Host mirror of the reference NEP force evaluator and integrator interfaces.

``NEP`` follows gpumd_compat::NEP (src/gpumd_compat/nep.cuh:144-238):
construct once from nep.txt, then ``compute(box, type, position, potential,
force, virial)`` on device arrays (the Potential::compute signature,
potential.cuh:82-88), or ``compute_domain`` with force / dependency centre
ranges (nep.cuh:178-189).  Arrays are torch tensors resident on the GPU in the
reference's structure-of-arrays layout; all work runs in libdmgx.so.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import DmgxError, check, lib

TIME_UNIT_CONVERSION = 1.018051e1     # fs per natural time unit (common.cuh:61)
PRESSURE_UNIT_CONVERSION = 1.602177e2  # GPa per eV/A^3 (common.cuh:60)
K_B = 8.617343e-5                      # eV/K (common.cuh:55)


@dataclass
class Box:
    """Simulation box; ``h`` has the cell vectors a, b, c as columns (the
    reference's cpu_h[0..8], box.cuh:53-70)."""

    h: np.ndarray
    pbc: tuple = (1, 1, 1)

    @staticmethod
    def from_lattice(a, b, c, pbc=(1, 1, 1)) -> "Box":
        return Box(np.column_stack([a, b, c]).astype(np.float64), tuple(int(p) for p in pbc))

    @staticmethod
    def orthogonal(lx, ly, lz, pbc=(1, 1, 1)) -> "Box":
        return Box(np.diag([lx, ly, lz]).astype(np.float64), tuple(int(p) for p in pbc))

    def to_c(self) -> _lib.Box:
        b = _lib.Box()
        flat = np.ascontiguousarray(self.h, dtype=np.float64).reshape(9)
        for k in range(9):
            b.h[k] = float(flat[k])
        for k in range(3):
            b.pbc[k] = int(self.pbc[k])
        return b

    @property
    def volume(self) -> float:
        return abs(float(np.linalg.det(self.h)))


def _ptr(t, dtype, numel, name):
    """Device pointer of a contiguous CUDA tensor; raises otherwise."""
    import torch
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor on the GPU")
    if not t.is_cuda:
        raise ValueError(f"{name} must live in device memory (got {t.device})")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.numel() < numel:
        raise ValueError(f"{name} holds {t.numel()} elements, needs {numel}")
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(int(getattr(stream, "cuda_stream", stream) or 0))


class NEP:
    """NEP4/NEP5 (+ZBL) evaluator on the GPU (the reference's NEP class)."""

    def __init__(self, potential_file: str):
        h = ctypes.c_void_p()
        check(lib().dmgx_nep_create(potential_file.encode(), ctypes.byref(h)))
        self._h = h
        info = _lib.NepInfo()
        check(lib().dmgx_nep_get_info(h, ctypes.byref(info)))
        self.info = info
        self.num_types = info.num_types
        self.dim = info.dim
        self.rc = info.rc_radial_max

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib._lib is not None:
            _lib._lib.dmgx_nep_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def reserve(self, capacity: int) -> None:
        check(lib().dmgx_nep_reserve(self._h, int(capacity)))

    def compute(self, box: Box, type, position, potential, force, virial, *, accumulate=False,
                ann_bf16=False, ann_valu=False, ann_bf16_split=False, force_rebuild=False, stream=None) -> None:
        """Potential::compute: every atom is a force and dependency centre."""
        N = type.numel()
        self.compute_domain(box, N, type, position, potential, force, virial, n1=0, n2=N, nd1=0,
                            nd2=N, accumulate=accumulate, ann_bf16=ann_bf16, ann_valu=ann_valu, ann_bf16_split=ann_bf16_split,
                            force_rebuild=force_rebuild, stream=stream)

    def compute_domain(self, box: Box, num_atoms: int, type, position, potential, force, virial,
                       global_id=None, *, n1=0, n2=None, nd1=0, nd2=None, sort_by_gid=False,
                       accumulate=False, ann_bf16=False, ann_valu=False, ann_bf16_split=False, force_rebuild=False,
                       stream=None) -> None:
        """NEP::compute_domain: [n1,n2) force centres, [nd1,nd2) dependency
        centres, every slot of [0,num_atoms) a neighbour candidate.  The ANN
        runs in FP32 on the matrix cores where the model allows (single type,
        <= 64 neurons); ann_valu keeps it on the vector ALUs, ann_bf16_split splits
        every operand into three bfloat16 parts (six part products, FP32
        accuracy), ann_bf16 uses
        bfloat16 operands."""
        import torch
        N = int(num_atoms)
        n2 = N if n2 is None else n2
        nd2 = N if nd2 is None else nd2
        flags = ((_lib.ACCUMULATE if accumulate else 0) | (_lib.ANN_BF16 if ann_bf16 else 0) |
                 (_lib.ANN_BF16_SPLIT if ann_bf16_split else 0) |
                 (_lib.FORCE_REBUILD if force_rebuild else 0) |
                 (_lib.SORT_BY_GID if sort_by_gid else 0) | (_lib.ANN_VALU if ann_valu else 0))
        b = box.to_c()
        check(lib().dmgx_nep_compute(
            self._h, ctypes.byref(b), N, n1, n2, nd1, nd2,
            _ptr(type, torch.int32, N, "type"), _ptr(position, torch.float64, 3 * N, "position"),
            _ptr(global_id, torch.int64, N, "global_id") if global_id is not None else None,
            _ptr(potential, torch.float64, N, "potential"), _ptr(force, torch.float64, 3 * N, "force"),
            _ptr(virial, torch.float64, 9 * N, "virial"), flags, _stream(stream)))

    def notify_box_changed(self) -> None:
        check(lib().dmgx_nep_notify_box_changed(self._h))

    def neighbor_stats(self) -> dict:
        v = [ctypes.c_int() for _ in range(4)]
        check(lib().dmgx_nep_neighbor_stats(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("max_radial", "max_angular", "max_verlet", "rebuilds"), [x.value for x in v]))

    def neighbor_means(self) -> dict:
        """Mean radial / angular / Verlet row length over the last centres."""
        a, b, c = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        n = ctypes.c_int()
        check(lib().dmgx_nep_neighbor_totals(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                             ctypes.byref(n)))
        k = max(n.value, 1)
        return {"radial": a.value / k, "angular": b.value / k, "verlet": c.value / k, "centres": n.value}

    KERNELS = ("filter", "descriptor", "gather", "rebuild", "check", "integrate", "thermo")

    def set_profiling(self, on: bool, dispatch_only: bool = False) -> None:
        """on: HIP-event timing of every kernel group (events recorded on the
        stream around each); dispatch_only: the fused descriptor kernel only,
        with events attached to its dispatch (the stream is left as it is)."""
        check(lib().dmgx_nep_set_profiling(self._h, (2 if dispatch_only else 1) if on else 0))

    def profile_reset(self) -> None:
        check(lib().dmgx_nep_profile_reset(self._h))

    def profile(self) -> dict:
        """{kernel: (launches, total_ms)} from the HIP events recorded so far."""
        out = {}
        for k, name in enumerate(self.KERNELS):
            n, ms = ctypes.c_int(), ctypes.c_double()
            check(lib().dmgx_nep_profile_read(self._h, k, ctypes.byref(n), ctypes.byref(ms)))
            out[name] = (n.value, ms.value)
        return out

    def tuning(self) -> dict:
        """Tile shapes the per-step kernels use (and the autotuner's timings)."""
        t = _lib.Tuning()
        check(lib().dmgx_nep_get_tuning(self._h, ctypes.byref(t)))
        return _tuning_dict(t)

    def set_tiles(self, filter_lanes: int, gather_lanes: int) -> None:
        """Sets the tile shapes directly (filter 8/16/32, gather 4/8/16)."""
        check(lib().dmgx_nep_set_tiles(self._h, int(filter_lanes), int(gather_lanes)))

    def set_split(self, split: bool) -> None:
        """Typewise row split once per Verlet rebuild (True) or every step
        (False, the reference's filter pass); the next evaluation rebuilds."""
        check(lib().dmgx_nep_set_split(self._h, 1 if split else 0))

    def set_desc_lanes(self, lanes: int) -> None:
        """Lanes per atom of the descriptor kernel: 1, 4, or 0 (by system
        size: four below 65,536 centres)."""
        check(lib().dmgx_nep_set_desc_lanes(self._h, int(lanes)))

    def autotune(self, box: Box, type, position, potential, force, virial, *, ann_bf16=False,
                 stream=None) -> dict:
        """Times the filter and gather tile shapes on this system and keeps the
        fastest for every later evaluation; the outputs hold the system's
        values afterwards."""
        import torch
        N = type.numel()
        t = _lib.Tuning()
        b = box.to_c()
        check(lib().dmgx_nep_autotune(
            self._h, ctypes.byref(b), N, _ptr(type, torch.int32, N, "type"),
            _ptr(position, torch.float64, 3 * N, "position"), _ptr(potential, torch.float64, N, "potential"),
            _ptr(force, torch.float64, 3 * N, "force"), _ptr(virial, torch.float64, 9 * N, "virial"),
            _lib.ANN_BF16 if ann_bf16 else 0, _stream(stream), ctypes.byref(t)))
        return _tuning_dict(t)

    def _rows(self, fn, args, c1, c2, max_nb):
        n = c2 - c1
        counts = np.zeros(max(n, 1), np.int32)
        rows = np.zeros(max(n, 1) * max_nb, np.int32)
        check(fn(self._h, *args, c1, c2, max_nb, counts.ctypes.data_as(_lib.PI),
                 rows.ctypes.data_as(_lib.PI)))
        return [rows[k * max_nb:k * max_nb + counts[k]].copy() for k in range(n)]

    def verlet_rows(self, c1, c2, max_nb=1024):
        return self._rows(lib().dmgx_nep_copy_verlet_rows, (), c1, c2, max_nb)

    def radial_rows(self, c1, c2, max_nb=1024):
        return self._rows(lib().dmgx_nep_copy_pair_rows, (0,), c1, c2, max_nb)

    def angular_rows(self, c1, c2, max_nb=1024):
        return self._rows(lib().dmgx_nep_copy_pair_rows, (1,), c1, c2, max_nb)


ENSEMBLES = {"nve": 0, "nvt_ber": 1, "nvt_bdp": 2, "nvt_nhc": 3, "npt_ber": 4, "npt_scr": 5}


class MD:
    """Device-resident MD segment (run_segment, src/runtime.cu:680-849):
    velocity Verlet in FP64 with the run.in ensembles, forces from ``nep``."""

    def __init__(self, nep: NEP, box: Box, types, masses, positions, velocities):
        types = np.ascontiguousarray(types, np.int32)
        N = types.shape[0]
        m = np.ascontiguousarray(masses, np.float64)
        x = np.ascontiguousarray(positions, np.float64).reshape(3, N)
        v = np.ascontiguousarray(velocities, np.float64).reshape(3, N)
        b = box.to_c()
        h = ctypes.c_void_p()
        check(lib().dmgx_md_create(nep.handle, ctypes.byref(b), N, types.ctypes.data_as(_lib.PI),
                                   m.ctypes.data_as(_lib.PD), x.ctypes.data_as(_lib.PD),
                                   v.ctypes.data_as(_lib.PD), ctypes.byref(h)))
        self._h = h
        self.nep = nep
        self.N = N
        self.order = None  # stored index -> construction index, once reordered

    def set_step_graph(self, on: bool) -> None:
        """Replays one captured step per step (``dmgx_md_set_step_graph``)."""
        check(lib().dmgx_md_set_step_graph(self._h, 1 if on else 0))

    def reorder(self) -> np.ndarray:
        """Stores the atoms in cell order (``dmgx_md_reorder``); returns, for
        each stored atom, its index in the arrays given at construction."""
        o = np.zeros(self.N, np.int32)
        check(lib().dmgx_md_reorder(self._h, o.ctypes.data_as(_lib.PI)))
        self.order = o
        return o

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib._lib is not None:
            _lib._lib.dmgx_md_destroy(h)
            self._h = None

    def run(self, steps: int, dt_fs: float, ensemble: str = "nve", t_init=0.0, t_final=None,
            t_coupling=100.0, p_target_gpa=0.0, bulk_modulus_gpa=0.0, p_coupling=1000.0,
            seed=12345678, ann_bf16=False, ann_valu=False, ann_bf16_split=False) -> np.ndarray:
        e = _lib.Ensemble(ENSEMBLES[ensemble], float(t_init), float(t_init if t_final is None else t_final),
                          float(t_coupling), float(p_target_gpa), float(bulk_modulus_gpa),
                          float(p_coupling), int(seed))
        th = np.zeros(8)
        check(lib().dmgx_md_run(self._h, int(steps), float(dt_fs), ctypes.byref(e),
                                (_lib.ANN_BF16 if ann_bf16 else 0) | (_lib.ANN_VALU if ann_valu else 0) |
                                (_lib.ANN_BF16_SPLIT if ann_bf16_split else 0), th.ctypes.data_as(_lib.PD)))
        return th

    def arrays(self, input_order=False) -> dict:
        """Host copies of position, velocity, force, potential, virial, in
        storage order, or with ``input_order`` in the order given at
        construction (the same unless ``reorder`` was called)."""
        ptrs = [ctypes.c_void_p() for _ in range(7)]
        check(lib().dmgx_md_arrays(self._h, *[ctypes.byref(p) for p in ptrs]))
        out = {}
        for name, p, n in (("position", ptrs[0], 3), ("velocity", ptrs[1], 3), ("force", ptrs[2], 3),
                           ("potential", ptrs[3], 1), ("virial", ptrs[4], 9)):
            a = np.zeros(n * self.N)
            _copy_d2h(p, a)
            out[name] = a.reshape(n, self.N) if n > 1 else a
        if input_order and self.order is not None:
            for name, a in out.items():
                b = np.empty_like(a)
                b[..., self.order] = a
                out[name] = b
        return out

    def box(self) -> Box:
        b = _lib.Box()
        check(lib().dmgx_md_box(self._h, ctypes.byref(b)))
        return Box(np.array(b.h[:]).reshape(3, 3), tuple(b.pbc[:]))

    def autotune(self, ann_bf16=False, ann_bf16_split=False) -> dict:
        """NEP.autotune on the state's current positions."""
        t = _lib.Tuning()
        check(lib().dmgx_md_autotune(self._h, (_lib.ANN_BF16 if ann_bf16 else 0) |
                                     (_lib.ANN_BF16_SPLIT if ann_bf16_split else 0), ctypes.byref(t)))
        return _tuning_dict(t)


class Server:
    """Continuous-batching server of independent periodic systems (no
    reference counterpart): each request is a single point (steps=0) or an
    NVE run of `steps` steps, evaluated with one of the server's models
    (`model`, an index into the NEP handles given); requests share the GPU in
    one batch per model and each gets the result of its own periodic system.
    Requests are admitted while their device bytes fit `hbm_bytes` (0: half
    the device memory free at start).  The NEP handles are owned by the
    server while it lives."""

    def __init__(self, nep, max_atoms: int = 4_000_000, max_requests: int = 4096, ann_bf16=False,
                 ann_valu=False, margin: float = 2.0, graph: bool = True, hbm_bytes: int = 0, ann_bf16_split=False):
        neps = list(nep) if isinstance(nep, (list, tuple)) else [nep]
        cfg = _lib.ServerConfig(int(max_atoms), int(max_requests),
                                (_lib.ANN_BF16 if ann_bf16 else 0) | (_lib.ANN_VALU if ann_valu else 0) |
                                (_lib.ANN_BF16_SPLIT if ann_bf16_split else 0),
                                float(margin), 1 if graph else 0, int(hbm_bytes))
        arr = (ctypes.c_void_p * len(neps))(*[n.handle for n in neps])
        h = ctypes.c_void_p()
        check(lib().dmgx_server_create(arr, len(neps), ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.neps = neps
        self._n = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib._lib is not None:
            _lib._lib.dmgx_server_destroy(h)
            self._h = None

    def submit(self, box: Box, types, masses, positions, velocities=None, steps: int = 0,
               dt_fs: float = 1.0, model: int = 0) -> int:
        t = np.ascontiguousarray(types, np.int32)
        N = t.shape[0]
        m = np.ascontiguousarray(masses, np.float64)
        x = np.ascontiguousarray(positions, np.float64).reshape(3, N)
        v = None if velocities is None else np.ascontiguousarray(velocities, np.float64).reshape(3, N)
        b = box.to_c()
        rid = ctypes.c_longlong()
        check(lib().dmgx_server_submit(self._h, int(model), ctypes.byref(b), N, t.ctypes.data_as(_lib.PI),
                                       m.ctypes.data_as(_lib.PD), x.ctypes.data_as(_lib.PD),
                                       None if v is None else v.ctypes.data_as(_lib.PD), int(steps),
                                       float(dt_fs), ctypes.byref(rid)))
        self._n[rid.value] = N
        return rid.value

    def poll(self, rid: int, wait: bool = False):
        """None while pending; else potential[N], force[3,N], position[3,N],
        velocity[3,N], virial[9,N] and latency_ms."""
        N = self._n[rid]
        pe, f, x, v, w = np.zeros(N), np.zeros(3 * N), np.zeros(3 * N), np.zeros(3 * N), np.zeros(9 * N)
        lat = ctypes.c_double()
        done = ctypes.c_int()
        check(lib().dmgx_server_poll(self._h, ctypes.c_longlong(rid), 1 if wait else 0,
                                     pe.ctypes.data_as(_lib.PD), f.ctypes.data_as(_lib.PD),
                                     x.ctypes.data_as(_lib.PD), v.ctypes.data_as(_lib.PD),
                                     w.ctypes.data_as(_lib.PD), ctypes.byref(lat), ctypes.byref(done)))
        if not done.value:
            return None
        return {"potential": pe, "force": f.reshape(3, N), "position": x.reshape(3, N),
                "velocity": v.reshape(3, N), "virial": w.reshape(9, N), "latency_ms": lat.value}

    def drain(self) -> None:
        check(lib().dmgx_server_drain(self._h))

    def reset_stats(self) -> None:
        """Zeroes the statistics of a drained server (its buffers stay)."""
        check(lib().dmgx_server_reset_stats(self._h))

    def stats(self) -> dict:
        s = _lib.ServerStats()
        check(lib().dmgx_server_get_stats(self._h, ctypes.byref(s)))
        return {name: getattr(s, name) for name, _ in s._fields_}


_hip = None


def _tuning_dict(t) -> dict:
    return {"filter_lanes": t.filter_lanes, "gather_lanes": t.gather_lanes, "tuned": bool(t.tuned),
            "filter_ms": dict(zip((8, 16, 32), t.filter_ms[:])), "gather_ms": dict(zip((4, 8, 16), t.gather_ms[:])),
            "row_layout": t.row_layout, "row_capacity": t.row_capacity,
            "force_ms": dict(zip(("multiple_of_4", "power_of_2"), t.force_ms[:])),
            "overlap": t.overlap, "overlap_ms": dict(zip(("sequential", "overlapped"), t.overlap_ms[:])),
            "split": t.split, "split_ms": dict(zip(("filter", "split_rows"), t.split_ms[:])),
            "desc_lanes": t.desc_lanes, "desc_ms": dict(zip((1, 4), t.desc_ms[:]))}


def _copy_d2h(dev_ptr, host: np.ndarray) -> None:
    """hipMemcpy device -> host of a library-owned buffer (synchronous)."""
    global _hip
    if _hip is None:
        _hip = lib()  # resolves hipMemcpy in the HIP runtime libdmgx.so is bound to
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
    st = _hip.hipMemcpy(ctypes.c_void_p(host.ctypes.data), dev_ptr, ctypes.c_size_t(host.nbytes), 2)
    if st != 0:
        raise DmgxError(-3, f"hipMemcpy failed ({st})")
