"""Builds libdmgx.so (HIP kernels for gfx950 + C++ host + the C ABI) in-tree.

Usage: python dmg-md_amd/build.py   (or __graft_entry__.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdmgx.so")
SOURCES = ["kernels.hip", "neighbor.hip", "nep_model.cpp", "ensemble.cpp", "engine.cpp", "serve.cpp", "frontend.cpp", "noalloc.hip", "capi.cpp"]
ARCH = os.environ.get("DMGX_OFFLOAD_ARCH", "gfx950")


def command(out: str = LIB) -> list:
    hipcc = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    return [
        hipcc, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        # packed FP32 (v_pk_*_f32) stays off: scalar FP32 only
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument", "-Wl,-z,defs",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
        *[os.path.join(CSRC, s) for s in SOURCES],
        "-o", out,
    ]


CLI = os.path.join(HERE, "dmgx-run")


def cli_command(out: str = CLI) -> list:
    """The dmgx-run executable: a thin C++ caller of dmgx_run_directory."""
    return ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "dmgx_run.cpp"),
            "-L", HERE, "-ldmgx", "-Wl,-rpath,$ORIGIN", "-o", out]


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libdmgx.so")
    os.replace(tmp, LIB)
    cmd = cli_command(CLI + ".tmp")
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building dmgx-run")
    os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
