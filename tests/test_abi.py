"""The C ABI without a GPU: libfeec.so loads, exports every function that
include/feec.h declares, and its host-only entry points (meshers, synthetic
inputs, option defaults, shapes) agree with the oracle and the reference's
literal Kuhn lists. No compute call reaches the device here."""
import ctypes as C
import os
import re

import numpy as np

import formoniq_amd as fa
from formoniq_amd import _lib
from oracle import oracle as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "feec.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(feec_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers the header exactly
    assert sorted(_lib.SIGNATURES) == names


def test_status_strings_and_defaults():
    L = _lib.load()
    assert L.feec_status_string(0) == b"ok"
    assert L.feec_status_string(7) == b"not converged"
    opts = fa.solve_opts()
    assert opts.rtol == 1e-10 and opts.precision == fa.FP64_PC32 and opts.graph_iters % 2 == 0


def test_kuhn_grid_matches_oracle_and_reference(golden):
    for d, n in [(1, 4), (2, 3), (3, 2), (4, 2)]:
        cells, coords = fa.CartesianGrid(d, n).triangulate()
        oc, ox = o.kuhn(d, n)
        assert cells.shape == oc.shape
        assert sorted(map(tuple, np.sort(cells, 1))) == sorted(map(tuple, np.sort(oc, 1)))
        np.testing.assert_array_equal(coords, ox)
    cells, coords = fa.CartesianGrid(3, 1).triangulate()
    assert sorted(map(tuple, cells)) == sorted(map(tuple, golden["kuhn_unit_cube"]["cells"]))
    assert coords.tolist() == golden["kuhn_unit_cube"]["coords"]


def test_jitter_moves_interior_only():
    g = fa.CartesianGrid(3, 4)
    _, x = g.triangulate()
    y = g.jitter(x, 0.05, 2)
    inner = np.all((x > 0) & (x < 1), axis=1)
    assert (y[~inner] == x[~inner]).all()
    assert (np.abs(y[inner] - x[inner]) <= 0.05 / 4 + 1e-15).all()
    assert (y[inner] != x[inner]).any()
    # deterministic in the seed, restricted to the requested axes
    np.testing.assert_array_equal(y, g.jitter(x, 0.05, 2))
    z = g.jitter(x, 0.05, 2, axes=[1, 2])
    assert (z[:, 0] == x[:, 0]).all()


def test_normal_vector_statistics():
    v = fa.normal_vector(200001, 1)
    assert abs(v.mean()) < 0.01 and abs(v.std() - 1) < 0.01
    np.testing.assert_array_equal(v[:10], fa.normal_vector(10, 1))


def test_element_shapes():
    for d in range(1, 5):
        for k in range(d + 1):
            for kind in range(4):
                if k == 0 and kind != 0:
                    continue
                assert fa.element_shape(d, kind, k) == o.form_shape(d, kind, k)
