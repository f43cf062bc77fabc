"""ctypes front end of the FP64 CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker.  The product library never
imports, links or calls anything under oracle/.

The C restatement lives in nep_oracle.c; see nep_oracle.h for the map onto the
reference files and for how each function is pinned.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libnep_oracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, FP64, no FMA contraction)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        dp = P(ctypes.c_double)
        ip = P(ctypes.c_int)
        L.nepo_load.argtypes = [ctypes.c_char_p, P(ctypes.c_void_p), ctypes.c_char_p, ctypes.c_int]
        L.nepo_load.restype = ctypes.c_int
        L.nepo_free.argtypes = [ctypes.c_void_p]
        L.nepo_info.argtypes = [ctypes.c_void_p, ip, dp]
        L.nepo_set_param_rounding.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.nepo_compute.argtypes = [ctypes.c_void_p, dp, ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ip, dp, dp, dp, dp, dp, dp]
        L.nepo_compute.restype = ctypes.c_int
        L.nepo_neighbor_rows.argtypes = [dp, ip, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_int, ip, ip]
        L.nepo_neighbor_rows.restype = ctypes.c_int
        L.nepo_velocity_verlet.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, dp, dp, dp]
        L.nepo_wrap_positions.argtypes = [dp, ip, ctypes.c_int, dp]
        L.nepo_thermo.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, dp]
        L.nepo_constant_tables.argtypes = [dp, dp, dp, dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class OracleNEP:
    """FP64 NEP evaluator on the host (nep_oracle.c)."""

    def __init__(self, path: str, round_params_to_float: bool = False):
        err = ctypes.create_string_buffer(512)
        h = ctypes.c_void_p()
        rc = lib().nepo_load(path.encode(), ctypes.byref(h), err, 512)
        if rc:
            raise ValueError(f"oracle could not parse {path}: {err.value.decode()}")
        self._h = h
        ints = np.zeros(12, np.int32)
        reals = np.zeros(4, np.float64)
        lib().nepo_info(h, _i(ints), _d(reals))
        (self.version, self.num_types, self.n_max_radial, self.n_max_angular,
         self.basis_radial, self.basis_angular, self.l_max, self.num_L, self.dim,
         self.num_neurons, zbl, flex) = [int(v) for v in ints]
        self.zbl_enabled = bool(zbl)
        self.zbl_flexible = bool(flex)
        self.rc_radial_max, self.rc_angular_max = float(reals[0]), float(reals[1])
        if round_params_to_float:
            lib().nepo_set_param_rounding(h, 1)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.nepo_free(self._h)
            self._h = None

    def compute(self, h9, pbc, types, pos_soa, n1=None, n2=None, nd1=None, nd2=None,
                want_descriptor=False):
        """Returns dict(pe[N], force[3,N], virial[9,N]) (+ q, fp if asked)."""
        types = _i32(types)
        N = types.shape[0]
        n1 = 0 if n1 is None else n1
        n2 = N if n2 is None else n2
        nd1 = n1 if nd1 is None else nd1
        nd2 = n2 if nd2 is None else nd2
        pos = _f64(pos_soa).reshape(-1)
        assert pos.shape[0] == 3 * N
        h = _f64(h9).reshape(9)
        p = _i32(pbc).reshape(3)
        pe = np.zeros(N)
        f = np.zeros(3 * N)
        v = np.zeros(9 * N)
        q = np.zeros(N * self.dim) if want_descriptor else None
        fp = np.zeros(N * self.dim) if want_descriptor else None
        null = ctypes.POINTER(ctypes.c_double)()
        rc = lib().nepo_compute(self._h, _d(h), _i(p), N, n1, n2, nd1, nd2, _i(types), _d(pos),
                                _d(pe), _d(f), _d(v), _d(q) if q is not None else null,
                                _d(fp) if fp is not None else null)
        if rc:
            raise ValueError(f"oracle compute failed ({rc})")
        out = {"pe": pe, "force": f.reshape(3, N), "virial": v.reshape(9, N)}
        if want_descriptor:
            out["q"] = q.reshape(N, self.dim)
            out["fp"] = fp.reshape(N, self.dim)
        return out


def neighbor_rows(h9, pbc, pos_soa, rc, c1=0, c2=None, max_nb=512):
    """Sorted FP64-predicate neighbour rows, list of int arrays."""
    pos = _f64(pos_soa).reshape(-1)
    N = pos.shape[0] // 3
    c2 = N if c2 is None else c2
    counts = np.zeros(max(c2 - c1, 1), np.int32)
    rows = np.zeros(max(c2 - c1, 1) * max_nb, np.int32)
    rc_ = lib().nepo_neighbor_rows(_d(_f64(h9).reshape(9)), _i(_i32(pbc).reshape(3)), N, _d(pos),
                                   c1, c2, float(rc), max_nb, _i(counts), _i(rows))
    if rc_:
        raise ValueError(f"neighbor_rows failed ({rc_})")
    return [rows[k * max_nb:k * max_nb + counts[k]].copy() for k in range(c2 - c1)]


def velocity_verlet(first_half, dt, mass, pos_soa, vel_soa, force_soa):
    N = mass.shape[0]
    lib().nepo_velocity_verlet(int(first_half), N, float(dt), _d(mass), _d(pos_soa), _d(vel_soa),
                               _d(force_soa))


def wrap_positions(h9, pbc, pos_soa):
    N = pos_soa.shape[-1] if pos_soa.ndim == 2 else pos_soa.shape[0] // 3
    lib().nepo_wrap_positions(_d(_f64(h9).reshape(9)), _i(_i32(pbc).reshape(3)), N, _d(pos_soa))


def thermo(h9, mass, pe, vel_soa, virial_soa):
    out = np.zeros(8)
    N = mass.shape[0]
    lib().nepo_thermo(N, _d(_f64(h9).reshape(9)), _d(_f64(mass)), _d(_f64(pe)), _d(_f64(vel_soa)),
                      _d(_f64(virial_soa)), _d(out))
    return out


def constant_tables():
    c3 = np.zeros(80)
    c4 = np.zeros(5)
    c5 = np.zeros(1)
    zc = np.zeros(45 * 9)
    lib().nepo_constant_tables(_d(c3), _d(c4), _d(c5), _d(zc))
    return c3, c4, float(c5[0]), zc.reshape(45, 9)
