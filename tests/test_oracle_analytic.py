"""Pins the oracle's analytic-data functions (quadrature, source loads, Whitney
values, L2 projection and error, de Rham map) by the reference's own laws and
by closed forms, on the CPU.

- Grundmann-Moeller exactness on barycentric monomials and barycentric nodes
  (simplicial/src/atlas/quadrature.rs tests).
- The load of a Whitney interpolant is M_k c, and the L2 projection reproduces
  Whitney forms (fe.rs:142-166): the load and the mass are one bilinear form.
- fe_l2_error vanishes on an affine 0-form (fe.rs:171-187) and equals closed
  forms (|1|_(L2) = 1 and |x|_(L2) = sqrt(1/3) on the unit box).
- Constant forms lie in the Whitney space: their projection is the cochain of
  edge / face integrals computed by hand from the coordinates (pins the
  ambient pullback on jittered, non-symmetric cells).
- The de Rham map is a cochain map, R d = d R (project.rs:101-186 cases).
- The projection error of a smooth 1-form halves with h (first order).
Each planted mistake (oracle.MUTATION 41-46) breaks one of them.
"""
import itertools
import math

import numpy as np
import pytest

from oracle import oracle as o


def _mesh(d, n, jitter=0.0, seed=3):
    cells, coords = o.kuhn(d, n)
    if jitter:
        rng = np.random.default_rng(seed)
        inner = np.all((coords > 1e-12) & (coords < 1 - 1e-12), axis=1)
        coords = coords.copy()
        coords[inner] += jitter / n * (2 * rng.random((inner.sum(), d)) - 1)
    cx = o.Complex(d, cells)
    return cx, coords, cx.lengths_from_coords(coords)


@pytest.fixture
def mutation():
    def set_(m):
        o.MUTATION = m
    yield set_
    o.MUTATION = 0


# ---- quadrature ------------------------------------------------------------------
def _gm_exact(dim, s):
    b, w = o.gm_rule(dim, 2 * s + 1)
    worst = 0.0
    for degree in range(2 * s + 2):
        for alpha in itertools.product(range(degree + 1), repeat=dim + 1):
            if sum(alpha) != degree:
                continue
            got = float(np.sum(w * np.prod(b ** np.array(alpha), axis=1)))
            exact = math.factorial(dim) * math.prod(math.factorial(a) for a in alpha) / \
                math.factorial(dim + degree)
            worst = max(worst, abs(got - exact))
    return worst


def test_gm_rule_is_exact_on_barycentric_monomials():
    for dim in range(5):
        for s in range(4):
            assert _gm_exact(dim, s) < 1e-12, (dim, s)
            b, w = o.gm_rule(dim, 2 * s + 1)
            assert abs(w.sum() - 1) < 1e-12 and np.allclose(b.sum(axis=1), 1.0)
            assert len(w) == sum(math.comb(s - i + dim, dim) for i in range(s + 1))


def test_gm_degree_constructor_is_minimal():
    # degree 2s and 2s+1 share the rule of index s
    for dim in range(1, 4):
        for deg in range(8):
            b1, _ = o.gm_rule(dim, deg)
            b2, _ = o.gm_rule(dim, 2 * (deg // 2) + 1)
            assert b1.shape == b2.shape and np.array_equal(b1, b2)


# ---- loads against the mass ---------------------------------------------------------
def _load_of_whitney(d, n, k, jitter=0.2):
    cx, coords, L = _mesh(d, n, jitter)
    nk = cx.nsimplices(k)
    c = np.array([((i % 7) - 3.0) for i in range(nk)])
    vals = o.whitney_at_nodes(cx, k, c, 3)
    b = o.source_load(cx, L, k, 3, vals)
    Mc = o.assemble(cx, L, o.MASS, k).to_scipy() @ c
    return np.abs(b - Mc).max() / np.abs(Mc).max()


@pytest.mark.parametrize("d,n", [(1, 4), (2, 3), (3, 2), (4, 1)])
def test_load_of_whitney_interpolant_is_mass_times_cochain(d, n):
    for k in range(d + 1):
        assert _load_of_whitney(d, n, k) < 1e-12, (d, k)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_l2_projection_reproduces_whitney_forms(d):
    cx, coords, L = _mesh(d, 2)
    wc = o.WhitneyComplex(cx, L)
    for k in range(d + 1):
        nk = cx.nsimplices(k)
        c = np.array([((i % 7) - 3.0) for i in range(nk)])
        vals = o.whitney_at_nodes(cx, k, c, 3)
        x = o.l2_projection(wc, k, 3, vals)
        assert np.abs(x - c).max() < 1e-9, (d, k)


# ---- L2 error -----------------------------------------------------------------------
@pytest.mark.parametrize("d", [1, 2, 3])
def test_l2_error_vanishes_on_an_affine_zero_form(d):
    cx, coords, L = _mesh(d, 2)
    wc = o.WhitneyComplex(cx, L)
    pts3 = o.cell_quadrature_points(cx, coords, 3)
    exact = pts3[..., :1]  # x_0 as a 0-form
    x = o.l2_projection(wc, 0, 3, exact, "ambient", coords)
    assert o.fe_l2_error(cx, L, 0, x, exact, "ambient", coords) < 1e-9


@pytest.mark.parametrize("d", [1, 2, 3])
def test_l2_error_closed_forms(d):
    cx, coords, L = _mesh(d, 3, jitter=0.2)
    pts3 = o.cell_quadrature_points(cx, coords, 3)
    zero0 = np.zeros(cx.nsimplices(0))
    one = np.ones(pts3.shape[:2] + (1,))
    assert abs(o.fe_l2_error(cx, L, 0, zero0, one) - 1.0) < 1e-12
    assert abs(o.fe_l2_error(cx, L, 0, zero0, pts3[..., :1]) - math.sqrt(1 / 3)) < 1e-12
    # the top form dx^1..dx^d (ambient) has unit L2 norm on the unit box
    zerod = np.zeros(cx.nsimplices(d))
    assert abs(o.fe_l2_error(cx, L, d, zerod, one, "ambient", coords) - 1.0) < 1e-12


# ---- constant forms and the ambient pullback ---------------------------------------
def _edge_integrals(cx, coords, a):
    e = cx.skeleton(1)
    return (coords[e[:, 1]] - coords[e[:, 0]]) @ a


def _face_integrals_3d(cx, coords, w):
    # w = (w01, w02, w12) on dx0^dx1, dx0^dx2, dx1^dx2; integral = w(u, v) / 2! * 2!
    f = cx.skeleton(2)
    u = coords[f[:, 1]] - coords[f[:, 0]]
    v = coords[f[:, 2]] - coords[f[:, 0]]
    return 0.5 * (w[0] * (u[:, 0] * v[:, 1] - u[:, 1] * v[:, 0]) +
                  w[1] * (u[:, 0] * v[:, 2] - u[:, 2] * v[:, 0]) +
                  w[2] * (u[:, 1] * v[:, 2] - u[:, 2] * v[:, 1]))


@pytest.mark.parametrize("d", [2, 3])
def test_constant_one_form_projects_to_its_edge_integrals(d):
    cx, coords, L = _mesh(d, 3, jitter=0.3)
    wc = o.WhitneyComplex(cx, L)
    a = np.array([0.7, -1.3, 0.4][:d])
    pts = o.cell_quadrature_points(cx, coords, 1)
    vals = np.broadcast_to(a, pts.shape[:2] + (d,)).copy()
    x = o.l2_projection(wc, 1, 1, vals, "ambient", coords)
    assert np.abs(x - _edge_integrals(cx, coords, a)).max() < 1e-9
    # and the de Rham map gives the same cochain
    sp = o.simplex_quadrature_points(cx, coords, 1, 1)
    r = o.derham_map(cx, coords, 1, 1, np.broadcast_to(a, sp.shape[:2] + (d,)))
    assert np.abs(r - _edge_integrals(cx, coords, a)).max() < 1e-12


def test_constant_two_form_projects_to_its_face_integrals():
    cx, coords, L = _mesh(3, 2, jitter=0.3)
    wc = o.WhitneyComplex(cx, L)
    w = np.array([0.5, -0.25, 1.5])
    pts = o.cell_quadrature_points(cx, coords, 1)
    vals = np.broadcast_to(w, pts.shape[:2] + (3,)).copy()
    x = o.l2_projection(wc, 2, 1, vals, "ambient", coords)
    want = _face_integrals_3d(cx, coords, w)
    assert np.abs(x - want).max() < 1e-9
    sp = o.simplex_quadrature_points(cx, coords, 2, 1)
    r = o.derham_map(cx, coords, 2, 1, np.broadcast_to(w, sp.shape[:2] + (3,)))
    assert np.abs(r - want).max() < 1e-12


# ---- de Rham map: Stokes ----------------------------------------------------------
STOKES_CASES = [
    # (dim, grade, omega(points) -> comps, d omega(points) -> comps)
    (1, 0, lambda p: p[..., 0:1] ** 2, lambda p: 2 * p[..., 0:1]),
    (2, 0, lambda p: p[..., 0:1] * p[..., 1:2], lambda p: np.stack([p[..., 1], p[..., 0]], -1)),
    (2, 1, lambda p: np.stack([p[..., 1], 0 * p[..., 0]], -1), lambda p: -np.ones(p.shape[:-1] + (1,))),
    (3, 1, lambda p: np.stack([0 * p[..., 0], p[..., 2], 0 * p[..., 0]], -1),
     lambda p: np.stack([0 * p[..., 0], 0 * p[..., 0], -np.ones(p.shape[:-1])], -1)),
    (3, 2, lambda p: np.stack([0 * p[..., 0], 0 * p[..., 0], p[..., 0]], -1),
     lambda p: np.ones(p.shape[:-1] + (1,))),
]


def _stokes_defect(d, k, om, dom, n=2):
    cx, coords, L = _mesh(d, n, jitter=0.3)
    r0 = o.derham_map(cx, coords, k, 3, om(o.simplex_quadrature_points(cx, coords, k, 3)))
    r1 = o.derham_map(cx, coords, k + 1, 3, dom(o.simplex_quadrature_points(cx, coords, k + 1, 3)))
    D = cx.coboundary(k).to_scipy()
    return np.abs(D @ r0 - r1).max() / max(np.abs(r1).max(), 1e-300)


@pytest.mark.parametrize("case", range(len(STOKES_CASES)))
def test_derham_map_is_a_cochain_map(case):
    d, k, om, dom = STOKES_CASES[case]
    assert _stokes_defect(d, k, om, dom) < 1e-12


# ---- convergence ------------------------------------------------------------------
def test_l2_projection_error_is_first_order():
    def field(p):
        return np.stack([np.sin(np.pi * p[..., 1]), np.cos(np.pi * p[..., 0])], -1)
    errs = []
    for n in (4, 8, 16):
        cx, coords, L = _mesh(2, n)
        wc = o.WhitneyComplex(cx, L)
        x = o.l2_projection(wc, 1, 3, field(o.cell_quadrature_points(cx, coords, 3)), "ambient", coords)
        errs.append(o.fe_l2_error(cx, L, 1, x, field(o.cell_quadrature_points(cx, coords, 3)),
                                  "ambient", coords))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 1.7 < r1 < 2.3 and 1.7 < r2 < 2.3, errs


# ---- planted mistakes ---------------------------------------------------------------
def test_mutation_41_gm_sign_breaks_exactness(mutation):
    mutation(41)
    assert _gm_exact(3, 1) > 1e-3


@pytest.mark.parametrize("m", [42, 43])
def test_mutations_42_43_break_the_load_law(mutation, m):
    mutation(m)
    assert _load_of_whitney(3, 2, 2) > 1e-3


def test_mutation_44_breaks_constant_forms(mutation):
    mutation(44)
    cx, coords, L = _mesh(3, 2, jitter=0.3)
    wc = o.WhitneyComplex(cx, L)
    a = np.array([0.7, -1.3, 0.4])
    pts = o.cell_quadrature_points(cx, coords, 1)
    x = o.l2_projection(wc, 1, 1, np.broadcast_to(a, pts.shape[:2] + (3,)).copy(), "ambient", coords)
    assert np.abs(x - _edge_integrals(cx, coords, a)).max() > 1e-4


def test_mutation_45_breaks_stokes(mutation):
    mutation(45)
    d, k, om, dom = STOKES_CASES[4]
    assert _stokes_defect(d, k, om, dom) > 1e-3


def test_mutation_46_breaks_the_closed_form(mutation):
    mutation(46)
    cx, coords, L = _mesh(2, 3, jitter=0.2)
    pts3 = o.cell_quadrature_points(cx, coords, 3)
    one = np.ones(pts3.shape[:2] + (1,))
    assert abs(o.fe_l2_error(cx, L, 0, np.zeros(cx.nsimplices(0)), one) - 1.0) > 1e-3
