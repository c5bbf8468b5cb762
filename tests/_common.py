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


def _triples(rp, col, val):
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return {(int(r), int(c)): v for r, c, v in zip(rows, col, val)}


# Entries whose exact value is zero but whose floating-point evaluation in the
# reference's formula leaves roundoff (exactly-representable cells) are kept by
# the oracle and dropped by the GPU, whose element tables produce exact zeros.
# Only such entries may differ, and only below this fraction of the matrix max.
ROUNDOFF = 1e-14


def assert_csr_parity(gpu: "fa.Csr", ora: "o.Csr", tol=1e-12, what="", exact_pattern=False):
    """Pattern bit-exact (up to roundoff-residue entries unless exact_pattern),
    values within tol relative to the matrix max. Returns the number of
    roundoff-residue entries present in one pattern only."""
    assert gpu.shape == ora.shape, (what, gpu.shape, ora.shape)
    rp, col, val = gpu.host()
    scale = max(np.abs(ora.val).max() if ora.nnz else 0.0, 1e-300)
    if gpu.nnz == ora.nnz and (rp == ora.row_ptr).all() and (col == ora.col).all():
        if ora.nnz:
            assert rel_max(val, ora.val) <= tol, (what, rel_max(val, ora.val))
        return 0
    assert not exact_pattern, (what, "pattern differs", gpu.nnz, ora.nnz)
    g = _triples(rp, col, val)
    r = _triples(ora.row_ptr, ora.col, ora.val)
    only_g = set(g) - set(r)
    only_r = set(r) - set(g)
    for key in only_g:
        assert abs(g[key]) <= ROUNDOFF * scale, (what, "gpu-only entry", key, g[key])
    for key in only_r:
        assert abs(r[key]) <= ROUNDOFF * scale, (what, "oracle-only entry", key, r[key])
    common = set(g) & set(r)
    err = max((abs(g[k] - r[k]) for k in common), default=0.0) / scale
    assert err <= tol, (what, err)
    return len(only_g) + len(only_r)
