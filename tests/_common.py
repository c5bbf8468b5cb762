"""Shared input builders for the parity tests: the same mesh, geometry and
source go to the oracle and to libfeec."""
import numpy as np

import formoniq_amd as fa
from oracle import oracle as o


def kuhn_mesh(d, n, jitter=0.0, seed=0, minkowski=False, axes=None, shuffle_seed=None):
    """Kuhn box through libfeec's host mesher, jittered with its SplitMix64."""
    g = fa.CartesianGrid(d, n)
    cells, coords = g.triangulate()
    if jitter:
        coords = g.jitter(coords, jitter, seed, axes=axes)
    if shuffle_seed is not None:
        cells = cells[np.random.default_rng(shuffle_seed).permutation(len(cells))]
    return cells, coords


def both(d, n, jitter=0.0, seed=0, minkowski=False, axes=None, shuffle_seed=None):
    cells, coords = kuhn_mesh(d, n, jitter, seed, minkowski, axes, shuffle_seed)
    ocx = o.Complex(d, cells)
    oL = ocx.lengths_from_coords(coords, minkowski)
    gcx = fa.Complex(d, cells, nverts=coords.shape[0])
    geo = fa.MeshLengthsSq.from_coords(gcx, coords, minkowski)
    return ocx, oL, gcx, geo, cells, coords


def rel_max(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = np.abs(b).max() if b.size else 0.0
    return float(np.abs(a - b).max() / scale) if scale > 0 else float(np.abs(a - b).max())


def assert_csr_parity(gpu: "fa.Csr", ora: "o.Csr", tol=1e-12, what=""):
    """Pattern bit-exact, values within tol relative to the matrix max."""
    assert gpu.shape == ora.shape, (what, gpu.shape, ora.shape)
    assert gpu.nnz == ora.nnz, (what, gpu.nnz, ora.nnz)
    rp, col, val = gpu.host()
    np.testing.assert_array_equal(rp, ora.row_ptr, err_msg=what)
    np.testing.assert_array_equal(col, ora.col, err_msg=what)
    if ora.nnz:
        assert rel_max(val, ora.val) <= tol, (what, rel_max(val, ora.val))
