"""Shared input builders for the parity tests: the same mesh, geometry and
source go to the oracle and to libfeec."""
import numpy as np

import formoniq_amd as fa
from oracle import oracle as o


def kuhn_mesh(d, n, jitter=0.0, seed=0, minkowski=False, axes=None, shuffle_seed=None,
              perturb_all=0.0):
    """Kuhn box through libfeec's host mesher, jittered with its SplitMix64.
    perturb_all > 0 also moves boundary vertices (no exactly-representable cell)."""
    g = fa.CartesianGrid(d, n)
    cells, coords = g.triangulate()
    if jitter:
        coords = g.jitter(coords, jitter, seed, axes=axes)
    if perturb_all:
        rng = np.random.default_rng(seed + 77)
        coords = coords + perturb_all / n * (2 * rng.random(coords.shape) - 1)
    if shuffle_seed is not None:
        cells = cells[np.random.default_rng(shuffle_seed).permutation(len(cells))]
    return cells, coords


def both(d, n, jitter=0.0, seed=0, minkowski=False, axes=None, shuffle_seed=None, perturb_all=0.0):
    cells, coords = kuhn_mesh(d, n, jitter, seed, minkowski, axes, shuffle_seed, perturb_all)
    ocx = o.Complex(d, cells)
    oL = ocx.lengths_from_coords(coords, minkowski)
    gcx = fa.Complex(d, cells, nverts=coords.shape[0])
    geo = fa.MeshLengthsSq.from_coords(gcx, coords, minkowski)
    return ocx, oL, gcx, geo, cells, coords


def rel_max(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = np.abs(b).max() if b.size else 0.0
    return float(np.abs(a - b).max() / scale) if scale > 0 else float(np.abs(a - b).max())


def assert_csr_parity(gpu: "fa.Csr", ora: "o.Csr", tol=0.0, what=""):
    """Pattern bit-exact: row pointers and columns equal the oracle's, so every
    entry the reference keeps (any nonzero contribution, galerkin.rs:173) is
    kept and no other. Values equal the oracle's bit for bit (tol = 0.0, the
    default) or within tol relative to the matrix max."""
    assert gpu.shape == ora.shape, (what, gpu.shape, ora.shape)
    rp, col, val = gpu.host()
    assert gpu.nnz == ora.nnz, (what, "nnz", gpu.nnz, ora.nnz)
    assert (rp == ora.row_ptr).all(), (what, "row_ptr differs")
    assert (col == ora.col).all(), (what, "col differs")
    if not ora.nnz:
        return
    if tol == 0.0:
        bad = np.flatnonzero(val != ora.val)
        assert bad.size == 0, (what, "values differ", bad.size, bad[:5], val[bad[:5]], ora.val[bad[:5]])
    else:
        assert rel_max(val, ora.val) <= tol, (what, rel_max(val, ora.val))
