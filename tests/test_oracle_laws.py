"""The reference's own laws, applied to the oracle's KKT and Hodge-Dirac
assembly (CPU). These pin the parts of the oracle that no golden vector
covers: a wrong U_k convention, a dropped sign or a wrong mass factor in
hodge_dirac or assemble_mixed_kkt fails one of them, and each planted mistake
(oracle.MUTATION) is shown to.

  D^2 = -Delta, grade by grade          dirac.rs:684-729
  D^2 = Delta on a Minkowski box        dirac.rs:825-873
  the constant field is reproduced      dirac.rs:881-943
  sigma = codif u: M_{k-1} s = D^T M_k u  hodge.rs:77-86 (the KKT's first row)
  the Hodge-Laplace equation for u      elliptic.rs:54-114 with the global routes
  coherent orientation                  adjacent cells induce opposite signs
  scaling: M_k(l^2 s) = l^(d-2k) M_k(s) (homogeneity of vol and Lambda^k g^-1)
"""
import itertools
import math
import os
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as spla

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle as o  # noqa: E402

CAUSAL_TIME_SCALE = 0.7  # regge/src/mesher/cartesian.rs:21


@pytest.fixture
def mutation():
    """Set oracle.MUTATION for one test and restore it."""
    def set_(m):
        o.MUTATION = m
    yield set_
    o.MUTATION = 0


def unit_grid(d, n=2, minkowski=False):
    """CartesianGrid::new_unit(d, n) or CartesianGrid::minkowski(d, n)
    (cartesian.rs:185-208: time axis 0 scaled by 0.7, no lightlike edge)."""
    cells, coords = o.kuhn(d, n)
    coords = coords.copy()
    if minkowski:
        coords[:, 0] *= CAUSAL_TIME_SCALE
    cx = o.Complex(d, cells)
    L = cx.lengths_from_coords(coords, minkowski)
    return cx, coords, L


def jittered(d, n, amp, seed):
    cells, coords = o.kuhn(d, n)
    rng = np.random.default_rng(seed)
    inter = np.all((coords > 1e-12) & (coords < 1 - 1e-12), axis=1)
    coords = coords.copy()
    coords[inter] += amp / n * (2 * rng.random((inter.sum(), d)) - 1)
    cx = o.Complex(d, cells)
    return cx, coords, cx.lengths_from_coords(coords)


def seed_field(offsets):
    """dirac.rs:611-624: ((7 i + 3 k + 1) mod 11) - 5 on grade k."""
    parts = []
    for k in range(len(offsets) - 1):
        i = np.arange(offsets[k + 1] - offsets[k])
        parts.append(((7 * i + 3 * k + 1) % 11) - 5.0)
    return np.concatenate(parts)


def dirac_square_residual(c: o.WhitneyComplex, sign):
    """max over grades of |D^2 u - sign Delta u| / max(|Delta u|, 1), with the
    strong Dirac action M^-1 A per grade and Delta_h u_k =
    M_k^-1 (dif_both(k+1) u_k + dif_test(k)^T M_{k-1}^-1 dif_test(k) u_k)."""
    offs, mb, op = o.hodge_dirac(c, sign)
    n = c.dim
    A = op.to_scipy()
    lu = [spla.splu(c.mass(k).to_scipy().tocsc()) for k in range(n + 1)]

    def apply(u):
        mv = A @ u
        return np.concatenate([lu[k].solve(mv[offs[k]:offs[k + 1]]) for k in range(n + 1)])

    u = seed_field(offs)
    d2u = apply(apply(u))
    worst = 0.0
    for k in range(n + 1):
        uk = u[offs[k]:offs[k + 1]]
        up = c.dif_both(k + 1).to_scipy() @ uk if k < n else np.zeros_like(uk)
        if k > 0:
            B = c.dif_test(k).to_scipy()
            dn = B.T @ lu[k - 1].solve(B @ uk)
        else:
            dn = np.zeros_like(uk)
        lap = lu[k].solve(up + dn)
        r = np.linalg.norm(d2u[offs[k]:offs[k + 1]] - sign * lap) / max(np.linalg.norm(lap), 1.0)
        worst = max(worst, r)
    return worst


@pytest.mark.parametrize("d", [1, 2, 3])
def test_dirac_squared_is_negative_hodge_laplacian(d):
    """dirac.rs:684-729 on the reference's unit grid, and on a jittered mesh."""
    for cx, _, L in (unit_grid(d), jittered(d, 3, 0.1, d)):
        assert dirac_square_residual(o.WhitneyComplex(cx, L), -1.0) <= 1e-9


@pytest.mark.parametrize("d", [1, 2, 3])
def test_selfadjoint_dirac_squares_to_hodge_laplacian_on_minkowski(d):
    """dirac.rs:825-873: the self-adjoint operator squares to +Delta on the
    causally generic Minkowski box (indefinite masses, LU solves)."""
    cx, _, L = unit_grid(d, minkowski=True)
    assert (L != 0).all()
    assert dirac_square_residual(o.WhitneyComplex(cx, L), 1.0) <= 1e-9


def constant_field(cx: o.Complex, coords):
    """The de Rham coefficients of the constant form with every component 1 on
    every grade (dirac.rs:896-912): the integral of sum_I dx^I over each
    simplex, (1/k!) sum_I det(E_I) with E the simplex's edge vectors from its
    first vertex and E_I their components I."""
    d = cx.dim
    parts = []
    for k in range(d + 1):
        simp = cx.skeleton(k)
        if k == 0:
            parts.append(np.ones(len(simp)))
            continue
        E = coords[simp[:, 1:]] - coords[simp[:, :1]]  # [N][k][d]
        tot = np.zeros(len(simp))
        for I in itertools.combinations(range(d), k):
            tot += np.linalg.det(E[:, :, list(I)])
        parts.append(tot / math.factorial(k))
    return np.concatenate(parts)


def constant_field_error(cx, coords, L, mass_term=1.0):
    """solve_dirac_source with load m M u and essential data u for the constant
    field u: the solution is u itself (dirac.rs:881-943)."""
    c = o.WhitneyComplex(cx, L)
    offs, mb, op = o.hodge_dirac(c, 1.0)
    exact = constant_field(cx, coords)
    M, A = mb.to_scipy(), op.to_scipy()
    S = (A + mass_term * M).tocsr()
    load = mass_term * (M @ exact)
    keep = np.concatenate([cx.boundary_flags(k) == 0 for k in range(cx.dim + 1)])
    rhs = (load - S @ exact)[keep]
    x = spla.spsolve(S[keep][:, keep].tocsc(), rhs)
    sol = exact.copy()
    sol[keep] += x
    return np.abs(sol - exact).max() / np.abs(exact).max()


@pytest.mark.parametrize("d", [1, 2, 3])
def test_dirac_source_reproduces_constant_field(d):
    """dirac.rs:881-943, Riemannian and Minkowski (the constant field
    interpolated on the time-scaled coordinates themselves)."""
    cx, coords, L = unit_grid(d)
    assert constant_field_error(cx, coords, L) <= 1e-9
    cx, coords, L = unit_grid(d, minkowski=True)
    assert constant_field_error(cx, coords, L) <= 1e-9


def kkt_law_residuals(c: o.WhitneyComplex, k, seed=0):
    """Solve the oracle's mixed KKT by SuperLU and return the relative
    residuals of (1) M_{k-1} sigma = D_{k-1}^T M_k u (hodge.rs:77-86) and (2)
    M_k D_{k-1} sigma + D_k^T M_{k+1} D_k u + M_k H p = E^T f, both with the
    global coboundaries D and the assembled masses (not the KKT's blocks)."""
    rng = np.random.default_rng(seed)
    h = c.harmonics(k)
    f = rng.standard_normal(c.cx.nsimplices(k))
    kkt, rhs, ns, nu = o.assemble_mixed_kkt(c, f, k, h)
    x = spla.spsolve(kkt.to_scipy().tocsc(), rhs)
    sig, u, p = x[:ns], x[ns:ns + nu], x[ns + nu:]
    Mk = c.mass(k).to_scipy()
    b = rhs[ns:ns + nu]
    r1 = 0.0
    lhs = Mk @ (h @ p) if h.shape[1] else np.zeros(nu)
    if k > 0:
        D = c.dif(k - 1).to_scipy()
        a = c.mass(k - 1).to_scipy() @ sig
        r1 = np.linalg.norm(a - D.T @ (Mk @ u)) / max(np.linalg.norm(a), 1e-300)
        lhs = lhs + Mk @ (D @ sig)
    if k < c.dim:
        Dk = c.dif(k).to_scipy()
        lhs = lhs + Dk.T @ (c.mass(k + 1).to_scipy() @ (Dk @ u))
    r2 = np.linalg.norm(lhs - b) / np.linalg.norm(b)
    return r1, r2


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("rel", [False, True])
def test_kkt_solution_obeys_codif_and_hodge_laplace(d, rel):
    cx, _, L = jittered(d, 3, 0.1, 7)
    c = o.WhitneyComplex(cx, L, relative=rel)
    for k in range(d + 1):
        r1, r2 = kkt_law_residuals(c, k)
        assert r1 <= 1e-10 and r2 <= 1e-10, (k, r1, r2)


@pytest.mark.parametrize("d", [2, 3, 4])
def test_orientation_is_coherent(d):
    """The oriented fundamental chain sum_c o_c c has boundary supported on the
    boundary facets: adjacent cells induce opposite signs on their shared
    facet (D_{d-1}^T o = 0 there), and each boundary facet gets +-1."""
    cx, _, _ = jittered(d, 2, 0.0, 0)
    orient = cx.orientation().astype(float)
    assert set(np.unique(orient)) <= {-1.0, 1.0}
    chain = cx.coboundary(d - 1).to_scipy().T @ orient
    bnd = cx.boundary_flags(d - 1) == 1
    assert (chain[~bnd] == 0).all()
    assert (np.abs(chain[bnd]) == 1).all()


@pytest.mark.parametrize("d", [1, 2, 3, 4])
def test_scaling_law(d):
    """lengths * l^2 gives vol * l^d and Lambda^k g^-1 * l^(-2k), so M_k and
    dif_both(k) (built on M_k) scale by l^(d-2k), dif_trial(k) and dif_test(k)
    likewise. With l^2 = 4 every step scales by a power of two: bit-exact."""
    rng = np.random.default_rng(d)
    cells, coords = o.kuhn(d, 1)
    cx = o.Complex(d, cells)
    L = cx.lengths_from_coords(coords + 0.05 * rng.standard_normal(coords.shape))
    s = L[cx.cell_faces(1)[0]]
    for k in range(d + 1):
        for kind in (o.MASS, o.DIF_TRIAL, o.DIF_TEST, o.DIF_BOTH):
            if k == 0 and kind != o.MASS:
                continue
            base = o.element_matrix(d, kind, k, s)
            assert np.array_equal(o.element_matrix(d, kind, k, 4.0 * s), base * 2.0 ** (d - 2 * k))
            np.testing.assert_allclose(o.element_matrix(d, kind, k, 3.0 * s), base * 3.0 ** ((d - 2 * k) / 2),
                                       rtol=1e-13, atol=1e-13 * np.abs(base).max())


# ---------------------------------------------------------------------------
# the laws have teeth: each planted mistake breaks one
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("m", [11, 12, 13])
def test_dirac_laws_catch_planted_mistakes(m, mutation):
    cx, coords, L = unit_grid(2)
    c = o.WhitneyComplex(cx, L)
    assert dirac_square_residual(c, -1.0) <= 1e-9
    mutation(m)
    assert dirac_square_residual(c, -1.0) > 1e-3, m


@pytest.mark.parametrize("m", [21, 22, 23])
def test_kkt_laws_catch_planted_mistakes(m, mutation):
    cx, _, L = jittered(2, 3, 0.1, 7)
    c = o.WhitneyComplex(cx, L)
    mutation(m)
    r1, r2 = kkt_law_residuals(c, 1)
    assert max(r1, r2) > 1e-3, (m, r1, r2)
