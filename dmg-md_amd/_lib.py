"""ctypes binding of libdmgx.so (the C ABI declared in include/dmgx.h).

The product path is the HIP library only: if libdmgx.so is missing or fails
to load, every entry point raises (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdmgx.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "dmgx.h")

_lib = None


class DmgxError(RuntimeError):
    """A DMGX_E* status returned by the library, with its message."""

    def __init__(self, code: int, message: str):
        super().__init__(f"dmgx error {code}: {message}")
        self.code = code


class Box(ctypes.Structure):
    _fields_ = [("h", ctypes.c_double * 9), ("pbc", ctypes.c_int * 3)]


class NepInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "version", "num_types", "n_max_radial", "n_max_angular", "basis_size_radial",
        "basis_size_angular", "l_max", "num_L", "dim", "num_neurons", "zbl_enabled",
        "zbl_flexible", "MN_radial", "MN_angular")] + [
        ("rc_radial_max", ctypes.c_double), ("rc_angular_max", ctypes.c_double),
        ("atomic_number", ctypes.c_int * 16)]


class Ensemble(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("t_init", ctypes.c_double), ("t_final", ctypes.c_double),
                ("t_coupling", ctypes.c_double), ("p_target_gpa", ctypes.c_double),
                ("bulk_modulus_gpa", ctypes.c_double), ("p_coupling", ctypes.c_double),
                ("seed", ctypes.c_uint32)]


class Nhc(ctypes.Structure):
    _fields_ = [("position", ctypes.c_double * 4), ("momentum", ctypes.c_double * 4),
                ("mass", ctypes.c_double * 4)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_atoms", ctypes.c_longlong), ("max_requests", ctypes.c_int), ("flags", ctypes.c_uint),
                ("margin", ctypes.c_double), ("use_graph", ctypes.c_int), ("hbm_bytes", ctypes.c_longlong)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_longlong) for n in ("submitted", "completed", "frames", "atom_frames")] + [
        (n, ctypes.c_double) for n in ("p50_ms", "p99_ms", "mean_ms", "max_ms", "elapsed_s", "frames_per_s")] + [
        (n, ctypes.c_longlong) for n in ("iterations", "relayouts", "max_slots")] + [
        (n, ctypes.c_double) for n in ("relayout_s", "iterate_s", "retire_s")] + [
        (n, ctypes.c_longlong) for n in ("hbm_budget", "max_bytes_in_flight", "reshells",
                                            "rows_carried")] + [("reshell_s", ctypes.c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [("filter_lanes", ctypes.c_int), ("gather_lanes", ctypes.c_int),
                ("filter_ms", ctypes.c_double * 3), ("gather_ms", ctypes.c_double * 3), ("tuned", ctypes.c_int),
                ("row_layout", ctypes.c_int), ("row_capacity", ctypes.c_int), ("force_ms", ctypes.c_double * 2),
                ("overlap", ctypes.c_int), ("overlap_ms", ctypes.c_double * 2),
                ("split", ctypes.c_int), ("split_ms", ctypes.c_double * 2),
                ("desc_lanes", ctypes.c_int), ("desc_ms", ctypes.c_double * 2)]


ACCUMULATE = 1
FORCE_REBUILD = 2
ANN_BF16 = 4
SORT_BY_GID = 8
ANN_VALU = 16
ANN_BF16_SPLIT = 32

VP = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
U = ctypes.c_uint
PI = ctypes.POINTER(ctypes.c_int)
PD = ctypes.POINTER(ctypes.c_double)

_SIGNATURES = {
    "dmgx_last_error": (ctypes.c_char_p, []),
    "dmgx_nep_create": (I, [ctypes.c_char_p, ctypes.POINTER(VP)]),
    "dmgx_nep_destroy": (None, [VP]),
    "dmgx_nep_get_info": (I, [VP, ctypes.POINTER(NepInfo)]),
    "dmgx_nep_reserve": (I, [VP, I]),
    "dmgx_nep_compute": (I, [VP, ctypes.POINTER(Box), I, I, I, I, I, VP, VP, VP, VP, VP, VP, U, VP]),
    "dmgx_nep_compute_noalloc": (I, [VP, ctypes.POINTER(Box), I, I, I, I, I, VP, VP, VP, VP, VP, VP, U, VP]),
    "dmgx_nep_notify_box_changed": (I, [VP]),
    "dmgx_nep_neighbor_stats": (I, [VP, PI, PI, PI, PI]),
    "dmgx_nep_neighbor_totals": (I, [VP] + [ctypes.POINTER(ctypes.c_longlong)] * 3 + [PI]),
    "dmgx_nep_set_profiling": (I, [VP, I]),
    "dmgx_nep_profile_read": (I, [VP, I, PI, PD]),
    "dmgx_nep_profile_reset": (I, [VP]),
    "dmgx_nep_get_tuning": (I, [VP, ctypes.POINTER(Tuning)]),
    "dmgx_nep_set_split": (I, [VP, I]),
    "dmgx_nep_set_desc_lanes": (I, [VP, I]),
    "dmgx_nep_set_tiles": (I, [VP, I, I]),
    "dmgx_nep_autotune": (I, [VP, ctypes.POINTER(Box), I, VP, VP, VP, VP, VP, U, VP, ctypes.POINTER(Tuning)]),
    "dmgx_nep_copy_verlet_rows": (I, [VP, I, I, I, PI, PI]),
    "dmgx_nep_copy_pair_rows": (I, [VP, I, I, I, I, PI, PI]),
    "dmgx_velocity_verlet": (I, [I, I, D, VP, VP, VP, VP, VP]),
    "dmgx_wrap_positions": (I, [ctypes.POINTER(Box), I, VP, VP]),
    "dmgx_thermo": (I, [ctypes.POINTER(Box), I, VP, VP, VP, VP, PD, VP]),
    "dmgx_scale_velocity": (I, [I, D, VP, VP]),
    "dmgx_scale_positions": (I, [I, D, VP, VP]),
    "dmgx_rng_create": (I, [ctypes.c_uint32, ctypes.POINTER(VP)]),
    "dmgx_rng_destroy": (None, [VP]),
    "dmgx_gasdev": (D, [VP]),
    "dmgx_gamdev": (D, [I, VP]),
    "dmgx_resamplekin_sumnoises": (D, [I, VP]),
    "dmgx_resamplekin": (D, [D, D, I, D, VP]),
    "dmgx_nhc_init": (I, [I, D, D, D, ctypes.POINTER(Nhc)]),
    "dmgx_nhc_half_step": (D, [ctypes.POINTER(Nhc), D, D, D, D]),
    "dmgx_npt_scr_scale": (D, [D, D, D, D, D, D]),
    "dmgx_md_create": (I, [VP, ctypes.POINTER(Box), I, PI, PD, PD, PD, ctypes.POINTER(VP)]),
    "dmgx_md_destroy": (None, [VP]),
    "dmgx_md_run": (I, [VP, I, D, ctypes.POINTER(Ensemble), U, PD]),
    "dmgx_md_arrays": (I, [VP] + [ctypes.POINTER(VP)] * 7),
    "dmgx_md_box": (I, [VP, ctypes.POINTER(Box)]),
    "dmgx_md_autotune": (I, [VP, U, ctypes.POINTER(Tuning)]),
    "dmgx_md_reorder": (I, [VP, PI]),
    "dmgx_md_set_step_graph": (I, [VP, I]),
    "dmgx_server_create": (I, [ctypes.POINTER(VP), I, ctypes.POINTER(ServerConfig), ctypes.POINTER(VP)]),
    "dmgx_server_destroy": (None, [VP]),
    "dmgx_server_submit": (I, [VP, I, ctypes.POINTER(Box), I, PI, PD, PD, PD, I, D,
                               ctypes.POINTER(ctypes.c_longlong)]),
    "dmgx_server_poll": (I, [VP, ctypes.c_longlong, I, PD, PD, PD, PD, PD, PD, PI]),
    "dmgx_server_drain": (I, [VP]),
    "dmgx_server_get_stats": (I, [VP, ctypes.POINTER(ServerStats)]),
    "dmgx_server_reset_stats": (I, [VP]),
    "dmgx_run_directory": (I, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]),
}


def header_functions(path: str = HEADER) -> list:
    """Names of every function declared in include/dmgx.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dmgx_\w+)\s*\(", text)))


def lib():
    """Loads libdmgx.so; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        # One HIP runtime per process: torch ships its own libamdhip64 (soname
        # libamdhip64.so.7).  Loading torch first makes libdmgx.so's
        # dependency resolve to that same instance instead of a second copy.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != 0:
        raise DmgxError(status, lib().dmgx_last_error().decode())
