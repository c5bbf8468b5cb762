"""Exact topology and harmonic bases of the library against the oracle: Betti
numbers and integral (co)cycles bit for bit, both harmonic readings to 1e-10
of their largest entry (the device CG to rtol 1e-12 against the oracle's
dense least squares), and the mixed Hodge-Laplace solve with a harmonic border
of rank 1 and 2 on meshes that are not balls (u, sigma and p to 1e-8)."""
import numpy as np
import pytest

import formoniq_amd as fa
from oracle import oracle as o

from _homology_meshes import FIXTURES, squared_lengths

pytestmark = pytest.mark.gpu


def _both(name, rel=False):
    dim, cells, geom = FIXTURES[name]()
    ocx = o.Complex(dim, cells)
    L = squared_lengths(ocx, geom)
    gcx = fa.Complex(dim, cells, nverts=int(cells.max()) + 1)
    geo = fa.MeshLengthsSq.from_lengths(gcx, L)
    return dim, ocx, o.WhitneyComplex(ocx, L, relative=rel), gcx, fa.WhitneyComplex(geo, relative=rel)


def _rel_max(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_betti_numbers_and_generators_match_the_oracle(name):
    dim, ocx, _, gcx, _ = _both(name)
    for rel in (False, True):
        bc = fa.RELATIVE if rel else fa.FULL
        for k in range(dim + 1):
            assert gcx.betti_number(k, bc) == o.betti_number(ocx, k, rel), (rel, k)
            z = gcx.cohomology_generators(k, bc)
            c = gcx.homology_generators(k, bc)
            assert z.tolist() == o.cohomology_generators(ocx, k, rel), (rel, k)
            assert c.tolist() == o.homology_generators(ocx, k, rel), (rel, k)


@pytest.mark.parametrize("name", ["annulus", "torus", "circle", "two_sphere", "tunnel", "two_disks", "mobius"])
def test_harmonic_bases_match_the_oracle(name):
    for rel in (False, True):
        dim, ocx, owc, gcx, gwc = _both(name, rel)
        for k in range(dim + 1):
            integral, ortho = o.harmonics_exact(owc, k)
            g_ortho, g_int = gwc.harmonics(k), gwc.harmonics_integral(k)
            assert g_ortho.shape == ortho.shape and g_int.shape == integral.shape, (rel, k)
            if ortho.size:
                assert _rel_max(g_ortho, ortho) <= 1e-10, (rel, k)
                assert _rel_max(g_int, integral) <= 1e-10, (rel, k)


@pytest.mark.parametrize("name,k", [("annulus", 1), ("torus", 1), ("torus", 0), ("two_disks", 0),
                                    ("tunnel", 1)])
def test_solve_source_with_a_harmonic_border(name, k):
    """elliptic.rs:132-182 where b_k = 1 or 2 on meshes that are not balls."""
    dim, ocx, owc, gcx, gwc = _both(name)
    f = np.cos(1.0 + np.arange(ocx.nsimplices(k)))
    s_o, u_o, p_o, rep_o = o.solve_source(owc, f, k, precond="direct")
    assert rep_o.converged
    s_g, u_g, p_g, rep_g = fa.solve_source(gwc, f, k)
    assert rep_g.converged and rep_g.residual <= 1e-10
    assert len(p_g) == len(p_o) == o.betti_number(ocx, k)
    scale = max(np.abs(u_o).max(), np.abs(p_o).max(), 1e-300)
    assert np.abs(u_g - u_o).max() <= 1e-8 * scale
    assert np.abs(p_g - p_o).max() <= 1e-8 * scale
    if k:
        assert np.abs(s_g - s_o).max() <= 1e-8 * max(np.abs(s_o).max(), scale)


def test_annulus_at_scale_obeys_the_harmonic_laws():
    """A 96x96 Kuhn square without its central 32x32 block (32,768 triangles,
    not a Kuhn box): b = (1, 1, 0) and the harmonic 1-form is closed,
    co-closed and M-normalized, from the library's own operators."""
    n, hole = 96, 32
    cells, coords = o.kuhn(2, n)
    lo, hi = (n - hole) // 2, (n + hole) // 2
    box = np.arange(len(cells)) // 2
    bx, by = box % n, box // n
    keep = ~((bx >= lo) & (bx < hi) & (by >= lo) & (by < hi))
    used, inv = np.unique(cells[keep], return_inverse=True)  # the hole's inner vertices go
    cells, coords = inv.reshape(-1, 3).astype(np.int64), coords[used]
    gcx = fa.Complex(2, cells, nverts=len(coords))
    assert [gcx.betti_number(k) for k in range(3)] == [1, 1, 0]
    assert [gcx.betti_number(k, fa.RELATIVE) for k in range(3)] == [0, 1, 1]
    geo = fa.MeshLengthsSq.from_coords(gcx, coords)
    wc = fa.WhitneyComplex(geo)
    h = wc.harmonics(1)
    assert h.shape == (gcx.nsimplices(1), 1)
    M = wc.mass(1).to_scipy()
    D1 = gcx.coboundary(1).to_scipy()
    D0 = gcx.coboundary(0).to_scipy()
    assert abs(h[:, 0] @ (M @ h[:, 0]) - 1.0) <= 1e-12
    assert np.abs(D1 @ h[:, 0]).max() <= 1e-12 * np.abs(h).max()
    assert np.abs(D0.T @ (M @ h[:, 0])).max() <= 1e-9 * np.abs(M @ h[:, 0]).max()
    # the period around the hole is 1 in the integral reading
    c = gcx.homology_generators(1)
    assert c.shape[0] == 1
    assert abs(wc.harmonics_integral(1)[:, 0] @ c[0] - 1.0) <= 1e-10


def test_more_harmonic_forms_than_the_border_holds_are_unsupported():
    """Nine disjoint triangles: b_0 = 9 exceeds the KKT border's 8 columns."""
    tri = np.array([[0, 1, 2]])
    cells = np.vstack([tri + 3 * i for i in range(9)])
    coords = np.vstack([np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]) + [2.0 * i, 0.0] for i in range(9)])
    gcx = fa.Complex(2, cells, nverts=27)
    assert gcx.betti_number(0) == 9
    wc = fa.WhitneyComplex(fa.MeshLengthsSq.from_coords(gcx, coords))
    with pytest.raises(fa.FeecError) as e:
        wc.harmonics(0)
    assert e.value.name == "FEEC_E_UNSUPPORTED"
    assert wc.harmonics(1).shape == (gcx.nsimplices(1), 0)


def test_4d_slab_with_a_hole_through_two_axes():
    """4D: the Kuhn 3^4 grid without the boxes (1, 1, z, w), a square annulus
    times a square: b = (1, 1, 0, 0, 0), relative (0, 0, 0, 1, 1) (Lefschetz
    duality), and the harmonic 1-form closed, co-closed and M-normalized with
    the library's own operators (the oracle's dense rational elimination is too
    slow at this size)."""
    cells, coords = o.kuhn(4, 3)
    box = np.arange(len(cells)) // 24
    keep = ~((box % 3 == 1) & ((box // 3) % 3 == 1))
    used, inv = np.unique(cells[keep], return_inverse=True)
    cells, coords = inv.reshape(-1, 5).astype(np.int64), coords[used]
    gcx = fa.Complex(4, cells, nverts=len(coords))
    chi = sum((-1) ** k * gcx.nsimplices(k) for k in range(5))
    assert chi == 0
    assert [gcx.betti_number(k) for k in range(5)] == [1, 1, 0, 0, 0]
    assert [gcx.betti_number(k, fa.RELATIVE) for k in range(5)] == [0, 0, 0, 1, 1]
    wc = fa.WhitneyComplex(fa.MeshLengthsSq.from_coords(gcx, coords))
    h = wc.harmonics(1)[:, 0]
    M = wc.mass(1).to_scipy()
    assert abs(h @ (M @ h) - 1.0) <= 1e-12
    assert np.abs(gcx.coboundary(1).to_scipy() @ h).max() <= 1e-12 * np.abs(h).max()
    assert np.abs(gcx.coboundary(0).to_scipy().T @ (M @ h)).max() <= 1e-9 * np.abs(M @ h).max()
    c = gcx.homology_generators(1)
    assert c.shape[0] == 1 and abs(wc.harmonics_integral(1)[:, 0] @ c[0] - 1.0) <= 1e-10
    # every cycle generator is a cycle: its boundary (D_0^T c) vanishes exactly
    assert not np.any(gcx.coboundary(0).to_scipy().T @ c[0])
