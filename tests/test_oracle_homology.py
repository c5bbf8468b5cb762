"""The oracle's exact (co)homology and harmonic bases, pinned by the
reference's own topology laws and known answers (simplicial/src/topology/
homology.rs tests, cohomology.rs tests, formoniq/src/harmonic.rs tests), by
sympy's independent rational RREF for the generator choice, and by planted
mistakes (oracle.MUTATION 51-54)."""
import math

import numpy as np
import pytest
import sympy

from oracle import oracle as o

from _homology_meshes import FIXTURES, annulus, kuhn_ball, squared_lengths, torus


def _cx(name):
    dim, cells, _ = FIXTURES[name]()
    return o.Complex(dim, cells)


def _wc(dim, cells, geom, rel):
    cx = o.Complex(dim, cells)
    return cx, o.WhitneyComplex(cx, squared_lengths(cx, geom), relative=rel)


# ---- Betti numbers ---------------------------------------------------------------------------
KNOWN = {  # (absolute, relative) Betti numbers
    "cube1": ([1, 0], [0, 1]),           # homology.rs cube_is_contractible, box_lefschetz_duality
    "cube2": ([1, 0, 0], [0, 0, 1]),
    "cube3": ([1, 0, 0, 0], [0, 0, 0, 1]),
    "two_sphere": ([1, 0, 1], [1, 0, 1]),  # closed: relative = absolute
    "annulus": ([1, 1, 0], [0, 1, 1]),    # b_1 = 1; relative by Lefschetz duality
    "circle": ([1, 1], [1, 1]),          # FlatQuotient::unit_torus(1, 3)
    "torus": ([1, 2, 1], [1, 2, 1]),      # unit_torus(2, 3): b_k = binom(2, k)
    "tunnel": ([1, 1, 0, 0], [0, 0, 1, 1]),
    "two_disks": ([2, 0, 0], [0, 0, 2]),
    "mobius": ([1, 1, 0], [0, 0, 0]),     # non-orientable: no relative top class over Q
}


@pytest.mark.parametrize("name", sorted(KNOWN))
def test_betti_numbers_known_answers(name):
    cx = _cx(name)
    absolute, relative = KNOWN[name]
    assert [o.betti_number(cx, k) for k in range(cx.dim + 1)] == absolute
    assert [o.betti_number(cx, k, True) for k in range(cx.dim + 1)] == relative
    # Euler-Poincare (homology.rs:54-66)
    chi = sum((-1) ** k * cx.nsimplices(k) for k in range(cx.dim + 1))
    assert sum((-1) ** k * b for k, b in enumerate(absolute)) == chi


def test_rank_mod_matches_sympy_rank():
    rng = np.random.default_rng(3)
    for _ in range(20):
        a = rng.integers(-2, 3, size=(rng.integers(1, 7), rng.integers(1, 7)))
        assert o.exact_rank(a) == sympy.Matrix(a.tolist()).rank()


# ---- generators ------------------------------------------------------------------------------
def _sympy_quotient(a, b):
    """ker A / im B by sympy: its nullspace basis (one vector per free column,
    a 1 there) kept where the rank over Q grows, each made primitive."""
    A = sympy.Matrix(a.tolist()) if a.size else sympy.zeros(a.shape[0], a.shape[1])
    if a.shape[0] == 0:
        kern = [sympy.eye(a.shape[1])[:, j] for j in range(a.shape[1])]
    else:
        kern = A.nullspace()
    cols = [sympy.Matrix(b[:, j].tolist()) for j in range(b.shape[1])]
    span = sympy.Matrix.hstack(*cols) if cols else sympy.zeros(a.shape[1], 0)
    r = span.rank() if cols else 0
    out = []
    for v in kern:
        grown = sympy.Matrix.hstack(span, v) if span.shape[1] else v
        if grown.rank() > r:
            span, r = grown, r + 1
            den = sympy.ilcm(*[sympy.fraction(x)[1] for x in v])
            ints = [int(x * den) for x in v]
            g = math.gcd(*ints)
            ints = [x // g for x in ints]
            if next(x for x in ints if x) < 0:
                ints = [-x for x in ints]
            out.append(ints)
    return out


@pytest.mark.parametrize("name", ["annulus", "torus", "two_sphere", "circle", "cube2", "two_disks", "mobius"])
@pytest.mark.parametrize("rel", [False, True])
def test_generators_match_sympy(name, rel):
    cx = _cx(name)
    for k in range(cx.dim + 1):
        A, B = o.integral_coboundary(cx, k, rel), o.integral_coboundary(cx, k - 1, rel)
        assert o.cohomology_generators(cx, k, rel) == _sympy_quotient(A, B), k
        A2, B2 = o.integral_coboundary(cx, k - 1, rel).T, o.integral_coboundary(cx, k, rel).T
        assert o.homology_generators(cx, k, rel) == _sympy_quotient(A2, B2), k


@pytest.mark.parametrize("name", ["annulus", "torus", "tunnel", "two_sphere", "two_disks"])
def test_generators_are_independent_cocycles(name):
    """cohomology.rs / homology.rs laws: one generator per Betti number, each
    closed exactly, primitive with its first entry positive, independent
    modulo the boundaries, and a nonsingular Kronecker pairing."""
    cx = _cx(name)
    for rel in (False, True):
        for k in range(cx.dim + 1):
            z = o.cohomology_generators(cx, k, rel)
            c = o.homology_generators(cx, k, rel)
            b = o.betti_number(cx, k, rel)
            assert len(z) == len(c) == b
            Dk, Dm = o.integral_coboundary(cx, k, rel), o.integral_coboundary(cx, k - 1, rel)
            for v in z:
                assert not np.any(Dk @ np.array(v)) if Dk.size else True
                assert math.gcd(*v) == 1 and next(x for x in v if x) > 0
            for v in c:
                assert not np.any(Dm.T @ np.array(v)) if Dm.size else True
            if b:
                Z = np.array(z).T
                assert o.exact_rank(np.hstack([Dm, Z])) == o.exact_rank(Dm) + b
                P = np.array(z) @ np.array(c).T
                assert abs(np.linalg.det(P)) > 0.5


# ---- harmonic bases --------------------------------------------------------------------------
def _harmonic_laws(wc, k, integral, ortho, cyc):
    M = wc.mass(k).todense()
    Dk = wc.dif(k).todense()
    if Dk.size:
        assert np.abs(Dk @ ortho).max() <= 1e-12                           # closed
    D = wc.dif(k - 1).todense() if k >= 1 else np.zeros((M.shape[0], 0))
    if D.size:
        assert np.abs(D.T @ M @ ortho).max() <= 1e-11                      # co-closed
    assert np.abs(ortho.T @ M @ ortho - np.eye(ortho.shape[1])).max() <= 1e-12
    periods = integral.T @ np.array(cyc, dtype=float).T
    assert np.abs(periods - np.eye(len(cyc))).max() <= 1e-12            # period-normalized
    # one space in two bases
    assert np.linalg.matrix_rank(np.hstack([integral, ortho]), tol=1e-9) == ortho.shape[1]


@pytest.mark.parametrize("name", ["annulus", "torus", "two_sphere", "tunnel", "mobius"])
def test_harmonic_bases_obey_the_laws(name):
    dim, cells, geom = FIXTURES[name]()
    for rel in (False, True):
        cx, wc = _wc(dim, cells, geom, rel)
        for k in range(dim + 1):
            integral, ortho = o.harmonics_exact(wc, k)
            assert ortho.shape == (wc.ndofs(k), o.betti_number(cx, k, rel))
            if ortho.shape[1]:
                _harmonic_laws(wc, k, integral, ortho, o.homology_generators(cx, k, rel))


@pytest.mark.parametrize("dim,n", [(1, 4), (2, 3), (3, 2)])
def test_exact_path_reduces_to_the_closed_forms_on_a_ball(dim, n):
    cells, coords = kuhn_ball(dim, n)
    cx = o.Complex(dim, cells)
    L = cx.lengths_from_coords(coords)
    for rel in (False, True):
        wc = o.WhitneyComplex(cx, L, relative=rel)
        for k in range(dim + 1):
            h_exact = wc.harmonics(k, "exact")
            h_closed = wc.harmonics(k, "closed")
            assert h_exact.shape == h_closed.shape, (rel, k)
            if h_exact.size:
                assert np.abs(h_exact - h_closed).max() <= 1e-12 * np.abs(h_closed).max()


def test_solve_source_on_an_annulus_is_orthogonal_to_the_harmonics():
    """elliptic.rs:132-182 with a harmonic border: u is M-orthogonal to the
    harmonic 1-form and the harmonic part of the source comes out in p."""
    dim, cells, geom = annulus()
    cx, wc = _wc(dim, cells, geom, False)
    h = wc.harmonics(1)
    assert h.shape[1] == 1
    M = wc.mass(1).todense()
    f = np.cos(np.arange(cx.nsimplices(1)))
    sig, u, p, rep = o.solve_source(wc, f, 1, precond="direct")
    assert rep.converged
    # the border row h^T M u = 0 and p = h^T f hold to the MINRES tolerance
    assert abs(h[:, 0] @ M @ u) <= 1e-9 * np.linalg.norm(f)
    assert abs(p[0] - h[:, 0] @ f) <= 1e-9 * np.linalg.norm(f)


# ---- planted mistakes ------------------------------------------------------------------------
@pytest.mark.parametrize("mutation", [51, 52, 53, 54])
def test_planted_mistakes_are_caught(mutation, monkeypatch):
    monkeypatch.setattr(o, "MUTATION", mutation)
    caught = False
    for name in ("annulus", "torus"):
        dim, cells, geom = FIXTURES[name]()
        for rel in (False, True):
            cx, wc = _wc(dim, cells, geom, rel)
            try:
                for k in range(dim + 1):
                    A, B = o.integral_coboundary(cx, k, rel), o.integral_coboundary(cx, k - 1, rel)
                    A2, B2 = B.T, A.T
                    if (o.cohomology_generators(cx, k, rel) != _sympy_quotient(A, B)
                            or o.homology_generators(cx, k, rel) != _sympy_quotient(A2, B2)):
                        caught = True
                    integral, ortho = o.harmonics_exact(wc, k)
                    if ortho.shape[1]:
                        _harmonic_laws(wc, k, integral, ortho, o.homology_generators(cx, k, rel))
            except (AssertionError, np.linalg.LinAlgError, ValueError):
                caught = True
    assert caught, mutation
