"""GPU parity of the analytic-data path (quadrature rules and nodes, SourceForm
loads, the Whitney interpolant at nodes, fe_l2_error, l2_projection and the de
Rham map) against the oracle, and its laws at the 3D 64^3 size.

Tolerances (FP64 throughout): node positions, Whitney values, loads, errors and
de Rham cochains agree with the oracle to 1e-12 relative to the largest entry
(the same operations in another association order); projections to 1e-9
(CG to a relative residual of 1e-12 on both sides)."""
import math

import numpy as np
import pytest

import formoniq_amd as fa
from oracle import oracle as o

from _common import both, rel_max

MESHES = [(1, 6, 0.2), (2, 4, 0.2), (3, 2, 0.2), (4, 1, 0.0)]


def test_quadrature_rules_equal_the_oracle():
    """Host code: runs without a GPU."""
    for dim in range(5):
        for deg in range(8):
            r = fa.SimplexQuadRule(dim, deg)
            b, w = o.gm_rule(dim, deg)
            np.testing.assert_array_equal(r.points, b)
            assert np.abs(r.weights - w).max() <= 4e-16
    with pytest.raises(fa.FeecError):
        fa.SimplexQuadRule(3, 8)
    with pytest.raises(fa.FeecError):
        fa.SimplexQuadRule(5, 1)


@pytest.mark.gpu
@pytest.mark.parametrize("d,n,amp", MESHES)
def test_nodes_loads_values_and_errors_match_the_oracle(d, n, amp):
    ocx, oL, gcx, geo, cells, coords = both(d, n, amp, seed=d)
    rng = np.random.default_rng(d)
    for degree in (1, 3, 5):
        pts = fa.cell_quadrature_points(geo, degree)
        assert rel_max(pts, o.cell_quadrature_points(ocx, coords, degree)) <= 1e-15
    for k in range(d + 1):
        nk = ocx.nsimplices(k)
        ncomp = math.comb(d, k)
        for degree in (1, 3):
            np_ = len(o.gm_rule(d, degree)[1])
            vals = rng.standard_normal((ocx.nsimplices(d), np_, ncomp))
            for frame in ("reference", "ambient"):
                g = fa.source_load(geo, k, vals, degree, frame)
                e = o.source_load(ocx, oL, k, degree, vals, frame, coords)
                assert rel_max(g, e) <= 1e-12, (k, degree, frame)
        c = rng.standard_normal(nk)
        gw = fa.whitney_at_nodes(geo, k, c, 3)
        assert rel_max(gw, o.whitney_at_nodes(ocx, k, c, 3)) <= 1e-12
        exact = rng.standard_normal(gw.shape)
        for frame in ("reference", "ambient"):
            ge = fa.fe_l2_error(c, exact, geo, k, frame)
            ee = o.fe_l2_error(ocx, oL, k, c, exact, frame, coords)
            assert abs(ge - ee) <= 1e-12 * ee, (k, frame)


@pytest.mark.gpu
@pytest.mark.parametrize("d,n,amp", MESHES)
def test_derham_map_matches_the_oracle(d, n, amp):
    ocx, oL, gcx, geo, cells, coords = both(d, n, amp, seed=d)
    rng = np.random.default_rng(10 + d)
    for k in range(d + 1):
        for degree in (1, 3):
            sp = fa.simplex_quadrature_points(geo, k, degree)
            assert rel_max(sp, o.simplex_quadrature_points(ocx, coords, k, degree)) <= 1e-15
            vals = rng.standard_normal(sp.shape[:2] + (math.comb(d, k),))
            g = fa.derham_map(geo, k, vals, degree)
            assert rel_max(g, o.derham_map(ocx, coords, k, degree, vals)) <= 1e-12, (k, degree)


@pytest.mark.gpu
@pytest.mark.parametrize("d,n", [(1, 8), (2, 4), (3, 2)])
def test_l2_projection_reproduces_whitney_forms(d, n):
    """fe.rs:142-166 on the device, against the oracle's projection."""
    ocx, oL, gcx, geo, cells, coords = both(d, n, 0.2, seed=4)
    owc = o.WhitneyComplex(ocx, oL)
    for k in range(d + 1):
        nk = ocx.nsimplices(k)
        c = np.array([((i % 7) - 3.0) for i in range(nk)])
        vals = fa.whitney_at_nodes(geo, k, c, 3)
        x, rep = fa.l2_projection(geo, k, vals, 3)
        assert rep.converged and np.abs(x - c).max() < 1e-9, (k, rep.as_dict())
        assert np.abs(x - o.l2_projection(owc, k, 3, vals)).max() < 1e-9
        # a reused mass gives the same bits
        xm, _ = fa.l2_projection(geo, k, vals, 3, mass=fa.WhitneyComplex(geo).mass(k))
        np.testing.assert_array_equal(xm, x)


@pytest.mark.gpu
def test_l2_projection_on_a_lorentzian_mesh_uses_minres():
    """An indefinite mass (Minkowski box): the projection still reproduces a
    Whitney form, as the reference's LU branch does (fe.rs:106-120)."""
    ocx, oL, gcx, geo, cells, coords = both(2, 4, 0.1, seed=2, minkowski=True)
    owc = o.WhitneyComplex(ocx, oL)
    for k in range(3):
        c = np.array([((i % 5) - 2.0) for i in range(ocx.nsimplices(k))])
        vals = fa.whitney_at_nodes(geo, k, c, 3)
        x, rep = fa.l2_projection(geo, k, vals, 3)
        assert rep.converged and np.abs(x - c).max() < 1e-8, (k, rep.as_dict())
        assert np.abs(x - o.l2_projection(owc, k, 3, vals)).max() < 1e-8
        xm, _ = fa.l2_projection(geo, k, vals, 3, mass=fa.WhitneyComplex(geo).mass(k))
        np.testing.assert_array_equal(xm, x)


@pytest.mark.gpu
def test_device_arrays_give_the_host_results():
    ocx, oL, gcx, geo, cells, coords = both(3, 3, 0.2, seed=5)
    rng = np.random.default_rng(7)
    vals = rng.standard_normal((ocx.nsimplices(3), 5, 3))
    dv = fa.DeviceArray.from_numpy(vals)
    h = fa.source_load(geo, 1, vals, 3, "ambient")
    g = fa.source_load(geo, 1, dv, 3, "ambient")
    np.testing.assert_array_equal(g.numpy(), h)
    x_h, _ = fa.l2_projection(geo, 1, vals, 3)
    x_d, _ = fa.l2_projection(geo, 1, dv, 3)
    np.testing.assert_array_equal(x_d.numpy(), x_h)
    c = rng.standard_normal(ocx.nsimplices(1))
    e_h = fa.fe_l2_error(c, vals, geo, 1)
    e_d = fa.fe_l2_error(fa.DeviceArray.from_numpy(c), dv, geo, 1)
    assert e_h == e_d
    w_d = fa.whitney_at_nodes(geo, 1, fa.DeviceArray.from_numpy(c), 3)
    np.testing.assert_array_equal(w_d.numpy(), fa.whitney_at_nodes(geo, 1, c, 3))
    pts = fa.cell_quadrature_points(geo, 3, device=True)
    np.testing.assert_array_equal(pts.numpy(), fa.cell_quadrature_points(geo, 3))
    sp = fa.derham_map(geo, 1, fa.DeviceArray.from_numpy(
        fa.simplex_quadrature_points(geo, 1, 1)), 1)
    np.testing.assert_array_equal(sp.numpy(), fa.derham_map(
        geo, 1, fa.simplex_quadrature_points(geo, 1, 1), 1))


@pytest.mark.gpu
def test_argument_errors():
    ocx, oL, gcx, geo, cells, coords = both(2, 2, 0.0)
    vals = np.zeros((ocx.nsimplices(2), 1, 2))
    with pytest.raises(ValueError):
        fa.source_load(geo, 1, vals, 1, "sideways")
    st = fa.load().feec_source_load(geo._h, 1, 1, 7, vals.ctypes.data, np.zeros(ocx.nsimplices(1)).ctypes.data, 0)
    assert st == 1  # FEEC_E_ARG for a frame outside feec_frame
    # an ambient frame needs coordinates
    geo_l = fa.MeshLengthsSq.from_lengths(gcx, oL)
    with pytest.raises(fa.FeecError):
        fa.source_load(geo_l, 1, vals, 1, "ambient")
    with pytest.raises(fa.FeecError):
        fa.source_load(geo, 3, vals, 1)
    with pytest.raises(fa.FeecError):
        fa.source_load(geo, 1, vals, 9)
    # a mass of the wrong grade is refused
    with pytest.raises(fa.FeecError):
        fa.l2_projection(geo, 1, vals, 1, mass=fa.WhitneyComplex(geo).mass(0))
    # a lengths-only geometry takes reference-frame fields
    assert np.array_equal(fa.source_load(geo_l, 1, vals, 1), fa.source_load(geo, 1, vals, 1))


# ---- 3D 64^3 (1.57 M cells): size-independent laws -----------------------------------
@pytest.fixture(scope="module")
def c2_geometry():
    from _common import kuhn_mesh
    cells, coords = kuhn_mesh(3, 64, 0.05, seed=2)
    gcx = fa.Complex(3, cells, nverts=coords.shape[0])
    geo = fa.MeshLengthsSq.from_coords(gcx, coords)
    return gcx, geo, coords


@pytest.mark.gpu
def test_c2_load_of_whitney_interpolant_is_the_galerkin_load(c2_geometry):
    """At full size: the quadrature load of W c equals M_1 c as feec_load_vector
    forms it (bit-pinned to the oracle's mass elsewhere)."""
    gcx, geo, coords = c2_geometry
    c = fa.normal_vector(gcx.nsimplices(1), 9)
    vals = fa.whitney_at_nodes(geo, 1, c, 3)
    b = fa.source_load(geo, 1, vals, 3)
    assert rel_max(b, fa.load_vector(geo, 1, c)) <= 1e-12


@pytest.mark.gpu
def test_c2_derham_map_is_a_cochain_map(c2_geometry):
    """Stokes at full size: R(d omega) = D_1 R(omega) for omega = z dy
    (d omega = -dy^dz), with the library's own coboundary."""
    gcx, geo, coords = c2_geometry
    p1 = fa.simplex_quadrature_points(geo, 1, 3)
    om = np.stack([0 * p1[..., 0], p1[..., 2], 0 * p1[..., 0]], -1)
    p2 = fa.simplex_quadrature_points(geo, 2, 1)
    dom = np.zeros(p2.shape[:2] + (3,))
    dom[..., 2] = -1.0
    r1 = fa.derham_map(geo, 1, om, 3)
    r2 = fa.derham_map(geo, 2, dom, 1)
    D = gcx.coboundary(1, fa.FULL).to_scipy()
    assert rel_max(D @ r1, r2) <= 1e-11


@pytest.mark.gpu
def test_projection_error_is_first_order_up_to_64():
    """examples/source.rs-style convergence on the device: the L2 projection
    error of a smooth 1-form halves with h from 16^3 to 64^3."""
    def field(p):
        return np.stack([np.sin(np.pi * p[..., 1]), np.cos(np.pi * p[..., 2]),
                         np.sin(np.pi * p[..., 0])], -1)
    from _common import kuhn_mesh
    errs = []
    for n in (16, 32, 64):
        cells, coords = kuhn_mesh(3, n)
        gcx = fa.Complex(3, cells, nverts=coords.shape[0])
        geo = fa.MeshLengthsSq.from_coords(gcx, coords)
        f3 = field(fa.cell_quadrature_points(geo, 3))
        x, rep = fa.l2_projection(geo, 1, f3, 3, "ambient")
        assert rep.converged
        errs.append(fa.fe_l2_error(x, f3, geo, 1, "ambient"))
    assert all(1.8 < errs[i] / errs[i + 1] < 2.2 for i in range(2)), errs
