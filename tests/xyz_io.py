"""Extended-XYZ reader for model.xyz inputs and the golden per-atom outputs.

Grammar as read by the reference's model parser (src/model_parser.cpp:279-416):
line 1 atom count, line 2 key=value pairs (Lattice, pbc, Properties, ...),
then one line per atom with the columns that Properties declares.
"""
from __future__ import annotations

import re
import shlex

import numpy as np

TIME_UNIT_CONVERSION = 1.018051e1  # fs per natural time unit (common.cuh:61)


def _kv(line: str) -> dict:
    out = {}
    for tok in shlex.split(line):
        if "=" in tok:
            k, v = tok.split("=", 1)
            out[k.strip().lower()] = v
    return out


def read_xyz(path: str, frame: int = 0) -> dict:
    with open(path) as fh:
        lines = fh.read().splitlines()
    pos = 0
    for _ in range(frame + 1):
        n = int(lines[pos].split()[0])
        header = lines[pos + 1]
        body = lines[pos + 2:pos + 2 + n]
        pos += 2 + n
    kv = _kv(header)
    lat = np.array([float(x) for x in kv["lattice"].split()])
    a, b, c = lat[0:3], lat[3:6], lat[6:9]
    h = np.column_stack([a, b, c])  # columns are the cell vectors
    pbc = [1, 1, 1]
    if "pbc" in kv:
        pbc = [1 if t.upper() == "T" else 0 for t in kv["pbc"].split()]
    props = kv.get("properties", "species:S:1:pos:R:3")
    fields = props.split(":")
    cols = {}
    col = 0
    for k in range(0, len(fields), 3):
        name, kind, width = fields[k].lower(), fields[k + 1], int(fields[k + 2])
        cols[name] = (col, width, kind)
        col += width
    rows = [ln.split() for ln in body]
    out = {"n": n, "h": h, "pbc": pbc, "header": kv}
    for name, (c0, w, kind) in cols.items():
        if kind == "S":
            out[name] = [r[c0] for r in rows]
        else:
            arr = np.array([[float(x) for x in r[c0:c0 + w]] for r in rows])
            out[name] = arr if w > 1 else arr[:, 0]
    if "energy" in kv:
        out["energy_total"] = float(kv["energy"])
    if "virial" in kv:
        out["virial_total"] = np.array([float(x) for x in kv["virial"].split()])
    return out


def read_thermo(path: str) -> list:
    """Rows of thermo.out as float arrays (comment lines skipped)."""
    rows = []
    with open(path) as fh:
        for ln in fh:
            if ln.lstrip().startswith("#") or not ln.strip():
                continue
            rows.append(np.array([float(x) for x in ln.split()]))
    return rows


def nep_symbols(nep_path: str) -> list:
    with open(nep_path) as fh:
        toks = fh.readline().split()
    return toks[2:2 + int(toks[1])]


def types_of(species, nep_path):
    sym = nep_symbols(nep_path)
    return np.array([sym.index(s) for s in species], dtype=np.int32)


# per-atom xyz virial column order -> internal SoA order xx yy zz xy xz yz yx zx zy
XYZ_VIRIAL_TO_SOA = [0, 4, 8, 1, 2, 5, 3, 6, 7]
