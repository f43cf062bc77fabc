"""Copies the reference's golden regression data (inputs + GPUMD 5.9.1 outputs)
into tests/golden/baseline as fixtures.

Source: tests/baseline/{inputs,goldens,manifest.json} of the reference tree
(these are data files: potentials, model.xyz/run.in inputs and per-atom
outputs at 17 significant digits).  Run once in a container that has the
reference mounted; the copied files are committed.
"""
import os
import shutil
import sys

SRC = "/root/reference/tests/baseline"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "baseline")


def main():
    if not os.path.isdir(SRC):
        sys.exit("reference baseline data not mounted")
    for sub in ("inputs", "goldens"):
        for root, _, files in os.walk(os.path.join(SRC, sub)):
            for f in files:
                s = os.path.join(root, f)
                d = os.path.join(DST, os.path.relpath(s, SRC))
                os.makedirs(os.path.dirname(d), exist_ok=True)
                shutil.copyfile(s, d)
    shutil.copyfile(os.path.join(SRC, "manifest.json"), os.path.join(DST, "manifest.json"))


if __name__ == "__main__":
    main()
