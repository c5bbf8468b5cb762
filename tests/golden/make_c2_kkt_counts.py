"""Writes tests/golden/c2_kkt_counts.json: the oracle's sparsity counts and
SHA-256 digests of the 3D 64^3 k=1 mixed KKT system (Kuhn cells, interior vertices jittered by
U(+-0.05 h) with seed 2, full complex), the configuration the benchmark
solves. The GPU test at full size compares libfeec's counts with these.

The counts come from the CPU oracle (oracle/liboracle.so, OpenMP): the
structural count (every element contribution, exact zeros included) and the
count the reference keeps (contributions pushed only when v != 0.0,
galerkin.rs:173), and digests of the CSR arrays (row_ptr as int64, col as
int32, val as float64, little-endian), so that the full-size pattern and
values can be compared bit for bit without the oracle. Run from the
repository root:
    python tests/golden/make_c2_kkt_counts.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import formoniq_amd as fa  # noqa: E402  (host mesher and jitter only)
from oracle import oracle as o  # noqa: E402


def digests(row_ptr, col, val):
    """SHA-256 of a CSR's arrays in a fixed byte layout."""
    return {"row_ptr": hashlib.sha256(np.ascontiguousarray(row_ptr, "<i8").tobytes()).hexdigest(),
            "col": hashlib.sha256(np.ascontiguousarray(col, "<i4").tobytes()).hexdigest(),
            "val": hashlib.sha256(np.ascontiguousarray(val, "<f8").tobytes()).hexdigest()}


def main(n=64):
    t0 = time.time()
    g = fa.CartesianGrid(3, n)
    cells, coords = g.triangulate()
    coords = g.jitter(coords, 0.05, 2)
    cx = o.Complex(3, cells)
    L = cx.lengths_from_coords(coords)
    c = o.WhitneyComplex(cx, L)
    h = c.harmonics(1)
    kkt, rhs, ns, nu = o.assemble_mixed_kkt(c, np.zeros(cx.nsimplices(1)), 1, h)
    blocks = {"mass0": c.mass(0).nnz, "dif_test1": c.dif_test(1).nnz, "dif_both2": c.dif_both(2).nnz,
              "mass1": c.mass(1).nnz, "dif_both1": c.dif_both(1).nnz}
    out = {"workload": f"3D Kuhn {n}^3, jitter 0.05h seed 2, k=1, full complex",
           "n_sigma": int(ns), "n_u": int(nu), "n_harmonic": int(h.shape[1]),
           "kkt_nnz": int(kkt.nnz), "kkt_nnz_structural": 60859905 if n == 64 else None,
           "block_nnz": blocks, "sha256": digests(kkt.row_ptr, kkt.col, kkt.val),
           "seconds": round(time.time() - t0, 1)}
    path = os.path.join(ROOT, "tests", "golden", "c2_kkt_counts.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 64)
