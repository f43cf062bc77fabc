"""ctypes front end of oracle/_ref/libnep_ref.so: the reference's own FP64
CPU NEP evaluator (nep_cpu.cpp) behind oracle/ref_driver.cpp.  Test
infrastructure and CPU-baseline timing only; built by oracle/ref.mk."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libnep_ref.so")
REF_SRC = "/root/reference/third_party/lammps-matpl-overlay/files/src/nep_cpu.cpp"
_lib = None


def build() -> bool:
    """Builds oracle/_ref when the reference tree is mounted; returns whether
    the library is available afterwards."""
    if os.path.exists(REF_SRC):
        subprocess.run(["make", "-s", "-f", os.path.join(_HERE, "ref.mk")],
                       cwd=os.path.dirname(_HERE), check=True)
    return os.path.exists(LIB_PATH)


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.refnep_create.argtypes = [ctypes.c_char_p]
        L.refnep_create.restype = ctypes.c_void_p
        L.refnep_destroy.argtypes = [ctypes.c_void_p]
        L.refnep_compute.argtypes = [ctypes.c_void_p, dp, ip, ctypes.c_int, ip, dp, dp, dp, dp]
        L.refnep_compute.restype = ctypes.c_int
        _lib = L
    return _lib


class RefNEP:
    def __init__(self, path: str):
        self._h = lib().refnep_create(path.encode())
        if not self._h:
            raise ValueError(f"reference evaluator could not load {path}")

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.refnep_destroy(self._h)
            self._h = None

    def compute(self, h9, pbc, types, pos_soa):
        types = np.ascontiguousarray(types, np.int32)
        N = types.shape[0]
        pos = np.ascontiguousarray(pos_soa, np.float64).reshape(-1)
        h = np.ascontiguousarray(h9, np.float64).reshape(9)
        p = np.ascontiguousarray(pbc, np.int32).reshape(3)
        pe = np.zeros(N)
        f = np.zeros(3 * N)
        v = np.zeros(9 * N)
        D = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        I = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        rc = lib().refnep_compute(self._h, D(h), I(p), N, I(types), D(pos), D(pe), D(f), D(v))
        if rc:
            raise ValueError("reference evaluator failed")
        return {"pe": pe, "force": f.reshape(3, N), "virial": v.reshape(9, N)}
