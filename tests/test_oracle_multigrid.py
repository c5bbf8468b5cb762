"""The oracle's multigrid and auxiliary-space preconditioner (the restatement
of formoniq/src/multigrid.rs, formoniq/src/hx.rs and iterative/src/multigrid.rs,
precond.rs, stationary.rs) pinned by what the mathematics fixes:

- the Kuhn-grid prolongation interpolates linear functions exactly and keeps
  the Dirichlet space (even grids);
- at grade 0 on an unjittered grid the Galerkin coarse operator P^T A P is the
  coarse grid's own hdif_gram(0) (multigrid.rs module note: "that the two
  agree at grade 0 is a test");
- the V-cycle is symmetric and contracts with an h-independent factor;
- the vector-nodal interpolation maps the constant coefficient fields e_I to
  the de Rham cochains of the constant forms dx^I (hx.rs:86-112);
- the auxiliary-space block preconditioner, plain or Chebyshev-accelerated,
  is symmetric positive definite, and the Lanczos estimate of its spectrum
  is a lower bound close enough to keep the polynomial positive;
- the reference's iteration law (elliptic.rs:338-367) holds for its exact
  block solves.

Each planted mistake (oracle.MUTATION 31-33) breaks one of them."""
import itertools
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle as o  # noqa: E402


@pytest.fixture
def mutation():
    def set_(m):
        o.MUTATION = m
    yield set_
    o.MUTATION = 0


def grid(d, n, amp=0.0, seed=0):
    cells, coords = o.kuhn(d, n)
    coords = coords.copy()
    if amp:
        rng = np.random.default_rng(seed)
        inter = np.all((coords > 1e-12) & (coords < 1 - 1e-12), axis=1)
        coords[inter] += amp / n * (2 * rng.random((inter.sum(), d)) - 1)
    cx = o.Complex(d, cells)
    return cx, coords, cx.lengths_from_coords(coords)


def linear_interpolation_error(d, g):
    """max |P (f on the coarse grid) - f on the fine grid| for an affine f;
    coarse vertex w sits at fine position 2 w (cells of width 2/g)."""
    _, fine = o.kuhn(d, g)
    gc = (g + 1) // 2
    _, coarse = o.kuhn(d, gc)
    coarse_pos = coarse * (2.0 * gc / g)
    a = np.array([0.3, -1.1, 0.7, 0.45][:d])
    P = o.kuhn_nodal_prolongation(d, g, False)
    return np.abs(P @ (coarse_pos @ a + 0.25) - (fine @ a + 0.25)).max()


@pytest.mark.parametrize("d,g", [(2, 4), (2, 5), (2, 7), (3, 4), (3, 5), (4, 3)])
def test_prolongation_interpolates_affine_functions(d, g):
    assert linear_interpolation_error(d, g) <= 1e-14


@pytest.mark.parametrize("d,g", [(2, 6), (3, 4)])
def test_prolongation_keeps_the_dirichlet_space(d, g):
    """Even g: a coarse function vanishing off the coarse interior prolongs to
    one vanishing on the fine boundary, and the interior prolongation is its
    restriction. (Odd g: the coarse grid overhangs the box by one fine cell
    and the interior prolongation spans a subspace of the fine Dirichlet
    space that is not a coarse finite-element space; it is still a Galerkin
    subspace.)"""
    _, fine = o.kuhn(d, g)
    gc = g // 2
    _, coarse = o.kuhn(d, gc)
    fi = np.rint(fine * g).astype(int)
    wi = np.rint(coarse * gc).astype(int)
    fin = np.all((fi > 0) & (fi < g), axis=1)
    cin = np.all((wi > 0) & (2 * wi <= g - 1), axis=1)
    f = np.random.default_rng(1).standard_normal(len(coarse)) * cin
    u = o.kuhn_nodal_prolongation(d, g, False) @ f
    assert np.abs(u[~fin]).max() == 0.0
    Pi = o.kuhn_nodal_prolongation(d, g, True)
    assert np.abs(Pi @ f[cin] - u[fin]).max() <= 1e-15


@pytest.mark.parametrize("d,n,rel", [(2, 8, False), (2, 8, True), (3, 4, False), (3, 4, True)])
def test_galerkin_coarse_operator_is_the_coarse_discretization(d, n, rel):
    cxf, _, Lf = grid(d, n)
    cxc, _, Lc = grid(d, n // 2)
    Af = o.WhitneyComplex(cxf, Lf, relative=rel).hdif_gram(0).to_scipy()
    Ac = o.WhitneyComplex(cxc, Lc, relative=rel).hdif_gram(0).to_scipy()
    G = o.kuhn_nodal_prolongation(d, n, rel).T @ Af @ o.kuhn_nodal_prolongation(d, n, rel)
    assert G.shape == Ac.shape
    assert abs(G - Ac).max() <= 1e-13 * abs(Ac).max()


def vcycle_contraction(n, rel=True, iters=40):
    cx, _, L = grid(2, n)
    A = o.WhitneyComplex(cx, L, relative=rel).hdif_gram(0).to_scipy().tocsr()
    V = o.NodalVCycle(A, 2, n, rel)
    x = np.random.default_rng(0).standard_normal(A.shape[0])
    lam = 0.0
    for _ in range(iters):
        y = x - V.apply(A @ x)
        lam = np.linalg.norm(y) / np.linalg.norm(x)
        x = y / np.linalg.norm(y)
    return lam, len(V.levels)


def test_vcycle_contraction_is_mesh_independent():
    """rho(I - V A) by power iteration on 2D relative grids of 1, 2 and 3
    multigrid levels: below 0.4 on each (an h-dependent smoother alone
    approaches 1 as h shrinks)."""
    out = [vcycle_contraction(n) for n in (16, 32, 64)]
    assert [lv for _, lv in out] == [1, 2, 3]
    for lam, _ in out:
        assert 0.05 < lam < 0.4, out


def vcycle_asymmetry(n=32):
    cx, _, L = grid(2, n, 0.1, 3)
    A = o.WhitneyComplex(cx, L, relative=True).hdif_gram(0).to_scipy().tocsr()
    V = o.NodalVCycle(A, 2, n, True)
    rng = np.random.default_rng(2)
    a, b = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    va, vb = V.apply(a), V.apply(b)
    return abs(va @ b - a @ vb) / abs(va @ b), min(a @ va, b @ vb)


def test_vcycle_is_symmetric_positive_definite():
    asym, quad = vcycle_asymmetry()
    assert asym <= 1e-12 and quad > 0


def constant_form_error(d, k):
    """max |Pi_k (e_I in every vertex) - integral of dx^I| over the k-simplices
    and the colex k-subsets I (sigma's edge vectors E from its first vertex:
    the integral is det(E_I) / k!)."""
    cx, coords, _ = grid(d, 3, 0.2, 1)
    Pi = o.vector_nodal_prolongation(cx, coords, k)
    n0 = cx.nsimplices(0)
    simp = cx.skeleton(k)
    E = coords[simp[:, 1:]] - coords[simp[:, :1]]
    err = 0.0
    for bi, I in enumerate(itertools.combinations(range(d), k)):
        x = np.zeros(Pi.shape[1])
        x[bi * n0:(bi + 1) * n0] = 1.0
        if k == 1:
            want = E[:, 0, I[0]]
        elif k == 2:
            want = 0.5 * (E[:, 0, I[0]] * E[:, 1, I[1]] - E[:, 0, I[1]] * E[:, 1, I[0]])
        else:
            want = np.linalg.det(E[:, :, list(I)]) / math.factorial(k)
        err = max(err, np.abs(Pi @ x - want).max())
    return err


@pytest.mark.parametrize("d,k", [(2, 1), (2, 2), (3, 1), (3, 2), (3, 3)])
def test_vector_nodal_interpolates_constant_forms(d, k):
    assert constant_form_error(d, k) <= 1e-14


@pytest.mark.parametrize("m,single", [(1, False), (4, False), (4, True)])
@pytest.mark.parametrize("d,n,k,rel", [(2, 8, 1, True), (2, 6, 2, False), (3, 3, 1, False)])
def test_hx_block_preconditioner_is_spd(d, n, k, rel, m, single):
    """Plain, Chebyshev-accelerated, and accelerated with the FP32 path's
    symmetric FP32 copy of hdif_gram in the Chebyshev steps."""
    cx, coords, L = grid(d, n, 0.1, 5)
    c = o.WhitneyComplex(cx, L, relative=rel)
    nh = c.harmonics(k).shape[1]
    prec = o.hx_block_preconditioner(c, coords, k, n, nh, cheb_m=m, single=single)
    N = c.ndofs(k - 1) + c.ndofs(k) + nh
    B = np.column_stack([prec(e) for e in np.eye(N)])
    assert np.abs(B - B.T).max() <= 1e-12 * np.abs(B).max()
    assert np.linalg.eigvalsh(0.5 * (B + B.T)).min() > 0


def minres_iters(n, precond):
    cx, coords, L = grid(2, n)
    c = o.WhitneyComplex(cx, L, relative=True)
    f = np.array([((i % 5) - 2.0) * 0.3 for i in range(cx.nsimplices(1))])
    if precond == "direct":
        *_, rep = o.solve_source(c, f, 1, precond="direct")
        return rep.iters
    kkt, rhs, _, _ = o.assemble_mixed_kkt(c, f, 1, c.harmonics(1))
    A = kkt.to_scipy()
    _, rep = o.minres_py(lambda v: A @ v, o.hx_block_preconditioner(c, coords, 1, n, 0), rhs, 1e-10, 5000)
    assert rep.converged
    return rep.iters


def test_reference_iteration_law_with_exact_blocks():
    """elliptic.rs:338-367: refinement 4 -> 8 adds at most 8 iterations."""
    coarse, fine = minres_iters(4, "direct"), minres_iters(8, "direct")
    assert fine <= coarse + 8, (coarse, fine)


def test_hx_iteration_count_grows_slowly():
    """The auxiliary-space blocks: 16^2 -> 32^2 adds at most 16 iterations."""
    coarse, fine = minres_iters(16, "hx"), minres_iters(32, "hx")
    assert fine <= coarse + 16, (coarse, fine)


def test_multigrid_laws_catch_planted_mistakes(mutation):
    mutation(31)
    assert linear_interpolation_error(2, 5) > 1e-3
    cxf, _, Lf = grid(2, 8)
    cxc, _, Lc = grid(2, 4)
    Af = o.WhitneyComplex(cxf, Lf).hdif_gram(0).to_scipy()
    Ac = o.WhitneyComplex(cxc, Lc).hdif_gram(0).to_scipy()
    P = o.kuhn_nodal_prolongation(2, 8, False)
    assert abs(P.T @ Af @ P - Ac).max() > 1e-3 * abs(Ac).max()
    mutation(32)
    assert vcycle_asymmetry()[0] > 1e-3
    mutation(33)
    assert constant_form_error(3, 1) > 1e-3 and constant_form_error(3, 2) > 1e-3


@pytest.mark.parametrize("d,n,k,rel", [(2, 8, 1, True), (3, 3, 1, False), (3, 4, 2, False)])
def test_lanczos_estimate_is_a_close_lower_bound(d, n, k, rel):
    """The 6-step Lanczos estimate of lambda_max(B_j hdif_gram(j)) is a Ritz
    value, so at most the true lambda_max (dense eigensolve of the
    A-symmetric B A), and above 0.8 of it. The Chebyshev polynomial of
    [b/4, b], b = 1.05 x estimate, is positive on (0, 1.25 b), so any
    estimate above true / 1.3125 keeps the block preconditioner SPD."""
    import scipy.linalg as sla
    cx, coords, L = grid(d, n, 0.1, 2)
    c = o.WhitneyComplex(cx, L, relative=rel)
    lam = o.hx_block_lambdas(c, coords, k, n)
    hu = o.HxPreconditioner(c, coords, k, n)
    hs = o.HxPreconditioner(c, coords, k - 1, n, nodal=hu.nodal)
    for b, (H, j) in enumerate([(hs, k - 1), (hu, k)]):
        A = c.hdif_gram(j).to_scipy().toarray()
        B = np.column_stack([H.apply(e) for e in np.eye(A.shape[0])])
        true = sla.eigh(A @ B @ A, A, eigvals_only=True)[-1]  # B A x = l x  <=>  A B A x = l A x
        assert true * 0.8 <= lam[b] <= true * (1 + 1e-9), (b, lam[b], true)


def test_chebyshev_hx_iteration_count_is_flat():
    """3D k = 1 with m = 4: 4^3 -> 8^3 -> 12^3 adds at most 3 iterations each."""
    counts = []
    for n in (4, 8, 12):
        cx, coords, L = grid(3, n, 0.05, 0)
        c = o.WhitneyComplex(cx, L)
        f = np.array([((i % 5) - 2.0) * 0.3 for i in range(cx.nsimplices(1))])
        kkt, rhs, _, _ = o.assemble_mixed_kkt(c, f, 1, c.harmonics(1))
        A = kkt.to_scipy()
        _, rep = o.minres_py(lambda v: A @ v, o.hx_block_preconditioner(c, coords, 1, n, 0, cheb_m=4),
                             rhs, 1e-10, 5000)
        counts.append(rep.iters)
    for coarse, fine in zip(counts, counts[1:]):
        assert fine <= coarse + 3, counts
