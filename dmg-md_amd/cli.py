"""Runs a run.in / model.xyz input directory (the reference's dmg-md
executable, src/dmgmd_main.cpp, on one GPU): commands are parsed and the MD
segments run by the library; thermo.out, dump_xyz files, restart.xyz and
neighbor.out are written into the directory.

Usage: python -m dmg-md_amd.cli [directory [run_file [model_file]]]
"""
from __future__ import annotations

import sys

from . import _lib


def run_directory(directory: str, run_file: str = "run.in", model_file: str = "model.xyz") -> None:
    _lib.check(_lib.lib().dmgx_run_directory(directory.encode(), run_file.encode(), model_file.encode()))


def main(argv=None) -> int:
    args = sys.argv[1:] if argv is None else argv
    if len(args) > 3:
        print(__doc__, file=sys.stderr)
        return 2
    try:
        run_directory(*(args or ["."]))
    except _lib.DmgxError as e:
        print(f"dmgx: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
