"""GPU parity of topology, element matrices and assembly against the oracle.

Everything is compared bit for bit: numbering, orientation signs, boundary
flags, cell metrics, element matrices, and the sparsity patterns and values of
the Galerkin matrices. The element kernel evaluates each entry with the
reference's operations in the reference's order (FP contraction off), so an
entry the reference's arithmetic leaves at a roundoff residue on an
exactly-representable cell is kept exactly where the oracle keeps it."""
import numpy as np
import pytest

import formoniq_amd as fa
from oracle import oracle as o

from _common import assert_csr_parity, both, rel_max

pytestmark = pytest.mark.gpu

MESHES = [(1, 7, 0.2), (2, 5, 0.1), (3, 3, 0.05), (4, 2, 0.02)]


@pytest.mark.parametrize("d,n,amp", MESHES)
def test_numbering_bit_exact(d, n, amp):
    ocx, oL, gcx, geo, cells, coords = both(d, n, amp, seed=d, shuffle_seed=11)
    for k in range(d + 1):
        assert gcx.nsimplices(k) == ocx.nsimplices(k)
        np.testing.assert_array_equal(gcx.skeleton(k), ocx.skeleton(k))
        np.testing.assert_array_equal(gcx.cell_faces(k), ocx.cell_faces(k))
        np.testing.assert_array_equal(gcx.boundary_flags(k), ocx.boundary_flags(k))
    np.testing.assert_array_equal(gcx.orientation(), ocx.orientation())
    for k in range(d):
        for bc in (fa.FULL, fa.RELATIVE):
            g = gcx.coboundary(k, bc)
            oc = ocx.coboundary(k)
            if bc == fa.RELATIVE:
                oc = oc.select(ocx.boundary_flags(k + 1) == 0, ocx.boundary_flags(k) == 0)
            assert_csr_parity(g, oc, 0.0, f"D_{k} bc={bc}")
    np.testing.assert_array_equal(geo.lengths(), oL)


def test_numbering_c1_config_size():
    ocx, oL, gcx, geo, _, _ = both(2, 32, 0.1, seed=0, shuffle_seed=3)
    for k in range(3):
        np.testing.assert_array_equal(gcx.skeleton(k), ocx.skeleton(k))
        np.testing.assert_array_equal(gcx.cell_faces(k), ocx.cell_faces(k))
    np.testing.assert_array_equal(geo.lengths(), oL)


@pytest.mark.parametrize("d,n,amp", MESHES + [(2, 3, 0.0), (3, 2, 0.0), (4, 1, 0.0)])
def test_element_matrices(d, n, amp):
    """Every block of every grade, bit-identical to the oracle's."""
    ocx, oL, gcx, geo, _, _ = both(d, n, amp, seed=d)
    wc = fa.WhitneyComplex(geo)
    for k in range(d + 1):
        for kind in range(4):
            if k == 0 and kind != fa.MASS:
                continue
            g = wc.element(kind, k)
            e = o.element_matrices(ocx, oL, kind, k)
            assert g.shape == e.shape
            bad = np.argwhere(g != e)
            assert bad.size == 0, (d, k, kind, len(bad), bad[:3], g[tuple(bad[0])], e[tuple(bad[0])])


def test_element_matrices_lorentzian():
    """4D Minkowski box, time axis unjittered: null and timelike edges present.
    Metric, inverse, determinant and all blocks bit-identical to the oracle."""
    ocx, oL, gcx, geo, _, _ = both(4, 2, 0.02, seed=3, minkowski=True, axes=[1, 2, 3])
    counts, riem = geo.causal_census()
    assert counts[0] > 0 and counts[1] > 0 and counts[2] > 0
    assert riem.sum() == 0  # no cell is Riemannian under diag(-1,1,1,1)
    assert counts.tolist() == [int((oL < 0).sum()), int((oL == 0).sum()), int((oL > 0).sum())]
    wc = fa.WhitneyComplex(geo)
    for k in range(5):
        for kind in range(4):
            if k == 0 and kind:
                continue
            g = wc.element(kind, k)
            e = o.element_matrices(ocx, oL, kind, k)
            assert np.array_equal(g, e), (k, kind, int((g != e).sum()))
    gm, gi, det = geo.cell_metrics()
    faces = ocx.cell_faces(1)
    for c in range(len(det)):
        og, ogi, odet = o.cell_metric(4, oL[faces[c]])
        assert np.array_equal(gm[c], og) and np.array_equal(gi[c], ogi) and det[c] == odet, c


@pytest.mark.parametrize("d,n,amp", MESHES + [(2, 4, 0.0), (3, 2, 0.0), (4, 2, 0.0), (3, 4, 0.05)])
def test_galerkin_matrices(d, n, amp):
    """Every Galerkin matrix of the full and the relative complex: pattern and
    values bit-identical. The unjittered grids and the boundary of a jittered
    one have exactly-representable cells."""
    ocx, oL, gcx, geo, _, _ = both(d, n, amp, seed=d + 10)
    for rel in (False, True):
        oc = o.WhitneyComplex(ocx, oL, relative=rel)
        gc = fa.WhitneyComplex(geo, relative=rel)
        for k in range(d + 1):
            assert_csr_parity(gc.mass(k), oc.mass(k), 0.0, f"mass {k} rel={rel}")
            if k >= 1:
                assert_csr_parity(gc.dif_trial(k), oc.dif_trial(k), 0.0, f"dif_trial {k} rel={rel}")
                assert_csr_parity(gc.dif_test(k), oc.dif_test(k), 0.0, f"dif_test {k} rel={rel}")
                assert_csr_parity(gc.dif_both(k), oc.dif_both(k), 0.0, f"dif_both {k} rel={rel}")
            assert_csr_parity(gc.hdif_gram(k), oc.hdif_gram(k), 0.0, f"hdif_gram {k} rel={rel}")


@pytest.mark.parametrize("d,n", [(2, 6), (3, 3), (4, 2)])
def test_galerkin_patterns_when_no_cell_is_exact(d, n):
    """Every vertex perturbed, boundary included: no element entry is zero by
    cancellation."""
    ocx, oL, gcx, geo, _, _ = both(d, n, 0.0, seed=d, perturb_all=0.1)
    for rel in (False, True):
        oc = o.WhitneyComplex(ocx, oL, relative=rel)
        gc = fa.WhitneyComplex(geo, relative=rel)
        for k in range(d + 1):
            assert_csr_parity(gc.mass(k), oc.mass(k), 0.0, f"mass {k}")
            if k >= 1:
                assert_csr_parity(gc.dif_test(k), oc.dif_test(k), 0.0, f"dif_test {k}")
                assert_csr_parity(gc.dif_both(k), oc.dif_both(k), 0.0, f"dif_both {k}")


def test_exact_zero_drop_unjittered():
    """On the unjittered 2x2 square the k=1 mass has exactly-zero element
    entries; the GPU pattern drops exactly the entries the oracle drops."""
    ocx, oL, gcx, geo, _, _ = both(2, 2)
    assert (o.element_matrices(ocx, oL, o.MASS, 1) == 0).any()
    assert_csr_parity(fa.WhitneyComplex(geo).mass(1), o.assemble(ocx, oL, o.MASS, 1), 0.0, "mass1")


def test_roundoff_residue_entries_kept():
    """A 3D 8^3 Kuhn mesh jittered like the 64^3 configuration (interior
    vertices only): the cells along the cube's edges are exactly
    representable, and the reference's arithmetic leaves entries there that
    are zero in exact arithmetic but a roundoff residue in floating point.
    The oracle keeps them (v != 0.0); so must the GPU, with the same bits."""
    ocx, oL, gcx, geo, _, _ = both(3, 8, 0.05, seed=2)
    oc = o.WhitneyComplex(ocx, oL)
    gc = fa.WhitneyComplex(geo)
    residues = 0
    for k, m in ((1, oc.mass(1)), (2, oc.dif_both(2))):
        residues += int((np.abs(m.val) <= 1e-14 * np.abs(m.val).max()).sum())
    assert residues > 0
    assert_csr_parity(gc.mass(1), oc.mass(1), 0.0, "mass1")
    assert_csr_parity(gc.dif_both(2), oc.dif_both(2), 0.0, "dif_both2")
    assert_csr_parity(gc.dif_test(1), oc.dif_test(1), 0.0, "dif_test1")


def test_fem3d_and_fdm_on_gpu(golden):
    from test_oracle import fem3d_galmat, kron_sum
    tets = golden["fem3d_tets"]["tets"]
    for n in (1, 2, 3, 5):
        _, _, gcx, geo, _, _ = both(3, n)
        A = fa.WhitneyComplex(geo).dif_both(1).to_scipy().toarray()
        np.testing.assert_allclose(A, fem3d_galmat(n, tets), atol=1e-12)
    for d in (2, 3, 4):
        n = 4 if d < 4 else 3
        g = fa.CartesianGrid(d, n)
        cells, coords = g.triangulate()
        coords = coords * n
        gcx = fa.Complex(d, cells, nverts=len(coords))
        geo = fa.MeshLengthsSq.from_coords(gcx, coords)
        wc = fa.WhitneyComplex(geo)
        A = wc.dif_both(1).to_scipy().toarray()
        b = wc.mass(0).to_scipy() @ np.ones(len(coords))
        A = A / b[:, None]
        inner = ~np.any((coords == 0) | (coords == n), axis=1)
        got = A[np.ix_(inner, inner)]
        one = np.diag(np.full(n - 1, 2)) - np.diag(np.ones(n - 2), 1) - np.diag(np.ones(n - 2), -1)
        assert np.abs(got - np.round(got)).max() < 1e-10
        assert (np.round(got) == kron_sum([one] * d)).all()


def test_degenerate_cell_is_an_error():
    cells, coords = fa.CartesianGrid(2, 2).triangulate()
    coords = coords.copy()
    coords[4] = coords[0]  # collapse the centre vertex onto a corner
    gcx = fa.Complex(2, cells, nverts=len(coords))
    geo = fa.MeshLengthsSq.from_coords(gcx, coords)
    with pytest.raises(fa.FeecError) as e:
        fa.WhitneyComplex(geo).mass(1)
    assert e.value.name == "FEEC_E_DEGENERATE_CELL"


def test_topology_errors():
    with pytest.raises(fa.FeecError) as e:
        fa.Complex(2, np.array([[0, 1, 2], [0, 1, 4]]), nverts=5)  # vertex 3 dangling
    assert e.value.name == "FEEC_E_TOPOLOGY"
    with pytest.raises(fa.FeecError) as e:
        fa.Complex(2, np.array([[0, 1, 2], [0, 1, 3], [0, 1, 4]]), nverts=5)  # facet in 3 cells
    assert e.value.name == "FEEC_E_TOPOLOGY"
