"""Writes the small golden fixtures used by the parity tests.

Run in the build container (the reference tree is only read here, never at
test time):

    python tests/golden/make_golden.py

1. ``dpotrf_tdg_mt8.json`` -- the task graph Nanos++ recorded for the
   reference's dpotrf at 8 x 8 tiles (visualize/TDG/dpotrf_tdg.dot): node ids
   in spawn order, task type (by the fill colour legend in that file) and the
   dependency edges.  Data only.
2. ``chol_small.npz`` -- inputs and outputs computed independently of both the
   oracle and the library: drand48 written out in pure Python from its POSIX
   definition, GENMAT_SPD's fill order (src/matutil/genmat.c:59-67, diagonal set
   to n), numpy/LAPACK Cholesky, and x = A^{-1} (A 1).
"""
import json
import re
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
DOT = Path("/root/reference/visualize/TDG/dpotrf_tdg.dot")
COLOR_TYPE = {"darkgreen": "potrf", "mediumpurple2": "trsm", "grey58": "gemm", "gray25": "syrk"}


def tdg_fixture():
    txt = DOT.read_text()
    nodes = {}
    for m in re.finditer(r'^\s*(\d+)\[style="filled", color="black", fillcolor="([a-z0-9]+)"\];', txt, re.M):
        nodes[int(m.group(1))] = COLOR_TYPE[m.group(2)]
    edges = []
    for m in re.finditer(r'^\s*(-?\d+) -> (-?\d+)\[', txt, re.M):
        a, b = int(m.group(1)), int(m.group(2))
        if a in nodes and b in nodes:
            edges.append([a, b])
    out = {"source": "visualize/TDG/dpotrf_tdg.dot", "mt": 8,
           "nodes": {str(k): v for k, v in sorted(nodes.items())}, "edges": sorted(edges)}
    (HERE / "dpotrf_tdg_mt8.json").write_text(json.dumps(out, indent=0))
    return out


def drand48_stream(seed, count):
    x = ((seed & 0xFFFFFFFF) << 16) | 0x330E
    out = np.empty(count)
    for i in range(count):
        x = (0x5DEECE66D * x + 0xB) & 0xFFFFFFFFFFFF
        out[i] = x / 2.0 ** 48
    return out


def genmat_spd_py(n, seed):
    r = drand48_stream(seed, n * (n + 1) // 2)
    A = np.zeros((n, n))
    q = 0
    for j in range(n):
        for i in range(j, n):
            d = float(n) if i == j else r[q]
            A[i, j] = A[j, i] = d
            q += 1
    return A


def chol_fixture():
    arrays = {}
    for n in (8, 64, 130):
        A = genmat_spd_py(n, 0)
        L = np.linalg.cholesky(A)
        b = A @ np.ones(n)
        arrays[f"A{n}"] = A
        arrays[f"L{n}"] = L
        arrays[f"b{n}"] = b
        arrays[f"x{n}"] = np.linalg.solve(A, b)
    arrays["drand48_seed7"] = drand48_stream(7, 16)
    np.savez_compressed(HERE / "chol_small.npz", **arrays)


if __name__ == "__main__":
    t = tdg_fixture()
    print("tdg nodes", len(t["nodes"]), "edges", len(t["edges"]))
    chol_fixture()
    print("chol_small.npz written")
