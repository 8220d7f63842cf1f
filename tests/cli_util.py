"""Helpers for the command-line tests: run the ``bart`` drop-in and move raw
BART-layout arrays through .hdr/.cfl files."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BART = os.path.join(ROOT, "bart_amd", "bin", "bart")


def run_bart(args, cwd, check=True):
    """Runs ``bart <args>`` in cwd; returns the CompletedProcess."""
    if not os.path.exists(BART):
        raise RuntimeError(f"{BART} is not built (run __graft_entry__.build())")
    p = subprocess.run([BART] + [str(a) for a in args], cwd=cwd, capture_output=True, text=True, timeout=300)
    if check and p.returncode != 0:
        raise AssertionError(f"bart {' '.join(map(str, args))} failed ({p.returncode}):\n{p.stderr}")
    return p


def write_raw(path, dims, data):
    """Writes the bytes of `data` (complex64, BART column-major order) with
    the given dims."""
    data = np.ascontiguousarray(data, dtype=np.complex64).reshape(-1)
    assert data.size == int(np.prod(dims))
    with open(path + ".hdr", "w") as f:
        f.write("# Dimensions\n" + " ".join(str(d) for d in dims) + " \n")
    data.tofile(path + ".cfl")


def read_raw(path):
    """Returns (dims, flat complex64 data) of a .hdr/.cfl pair."""
    with open(path + ".hdr") as f:
        lines = f.read().split("\n")
    dims = [int(v) for v in lines[lines.index("# Dimensions") + 1].split()]
    data = np.fromfile(path + ".cfl", dtype=np.complex64)
    assert data.size == int(np.prod(dims))
    return dims, data


def dims16(*d):
    return list(d) + [1] * (16 - len(d))
