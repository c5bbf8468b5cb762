"""The CPU oracle pinned against the reference's own known answers, against
closed forms that do not go through the oracle's formula, and against laws a
plausible mistake would break. No GPU.

Known answers (tests/golden/reference_known_answers.json) are the hand tables
and closed forms of formoniq's tests (crates/formoniq/tests/unit_elmat.rs,
crates/formoniq/src/operators.rs:921-997, crates/formoniq/tests/fdm.rs,
crates/formoniq/tests/fem3d.rs, the Kuhn lists of
crates/simplicial/src/mesher/grid.rs:115-138 and
crates/regge/src/mesher/cartesian.rs:242-301).
"""
import itertools
import math
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def refcell(d):
    return o.unit_lengths(d)


def jittered(d, n, amp, seed):
    cells, coords = o.kuhn(d, n)
    rng = np.random.default_rng(seed)
    inter = np.all((coords > 1e-12) & (coords < 1 - 1e-12), axis=1)
    coords = coords.copy()
    coords[inter] += amp / n * (2 * rng.random((inter.sum(), d)) - 1)
    return cells, coords


# ---------------------------------------------------------------------------
# the reference's known answers
# ---------------------------------------------------------------------------

def test_laplacian_refcell(golden):
    for d, expected in golden["laplacian_refcell"]["cases"].items():
        d = int(d)
        got = o.element_matrix(d, o.DIF_BOTH, 1, refcell(d))
        np.testing.assert_allclose(got, expected, rtol=1e-14, atol=1e-15)


def test_mass_refcell(golden):
    for d, expected in golden["mass_refcell"]["cases"].items():
        d = int(d)
        got = o.element_matrix(d, o.MASS, 0, refcell(d))
        np.testing.assert_allclose(got, expected, rtol=1e-14, atol=1e-15)


def test_hodge_mass_dim2_grade1_and_difs(golden):
    s = refcell(2)
    np.testing.assert_allclose(o.element_matrix(2, o.MASS, 1, s), golden["hodge_mass_dim2_grade1"]["value"],
                               rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(o.element_matrix(2, o.DIF_TRIAL, 1, s), golden["dif_trial_n2_k1"]["value"],
                               rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(o.element_matrix(2, o.DIF_TEST, 1, s), golden["dif_test_n2_k1"]["value"],
                               rtol=1e-14, atol=1e-15)


def test_mass0_is_vol_times_bary_gramian():
    # operators.rs:921-929: M_0 = vol * Q, Q = (1 + delta)/((d+1)(d+2))
    for d in range(1, 5):
        n = d + 1
        q = (np.ones((n, n)) + np.eye(n)) / (n * (n + 1))
        np.testing.assert_allclose(o.element_matrix(d, o.MASS, 0, refcell(d)), q / math.factorial(d),
                                   rtol=1e-14)


def test_kuhn_lists(golden):
    cells, coords = o.kuhn(3, 1)
    cx = o.Complex(3, cells)
    assert cx.cells().tolist() == golden["kuhn_unit_cube"]["cells"]
    assert coords.tolist() == golden["kuhn_unit_cube"]["coords"]
    cells, coords = o.kuhn(2, 2)
    cx = o.Complex(2, cells)
    assert cx.cells().tolist() == golden["kuhn_unit_square_2"]["cells"]
    assert coords.tolist() == golden["kuhn_unit_square_2"]["coords"]


def fem3d_galmat(n, tets):
    """crates/formoniq/tests/fem3d.rs:18-105: P1 stiffness by cross products."""
    nv1 = n + 1
    h = 1.0 / n
    x = np.zeros((nv1 ** 3, 3))
    for z in range(nv1):
        for y in range(nv1):
            for xx in range(nv1):
                x[xx + nv1 * (y + nv1 * z)] = (h * xx, h * y, h * z)
    tet_vol = h ** 3 / 6
    A = np.zeros((nv1 ** 3, nv1 ** 3))
    for zb, yb, xb in itertools.product(range(n), repeat=3):
        box = [xb + i + nv1 * (yb + j + nv1 * (zb + k)) for k in range(2) for j in range(2) for i in range(2)]
        for tet in tets:
            g = [box[i] for i in tet]
            ns = []
            for i in range(4):
                face = [x[v] for j, v in enumerate(g) if j != i]
                sign = 1.0 if i % 2 == 0 else -1.0
                ns.append(sign * np.cross(face[1] - face[0], face[2] - face[0]))
            el = np.array([[ni @ nj for nj in ns] for ni in ns]) / (36 * tet_vol)
            for a in range(4):
                for b in range(4):
                    A[g[a], g[b]] += el[a, b]
    return A


def test_fem3d(golden):
    tets = golden["fem3d_tets"]["tets"]
    for n in range(1, 5):
        cells, coords = o.kuhn(3, n)
        cx = o.Complex(3, cells)
        L = cx.lengths_from_coords(coords)
        feec = o.assemble(cx, L, o.DIF_BOTH, 1).todense()
        np.testing.assert_allclose(feec, fem3d_galmat(n, tets), atol=1e-12)


def kron_sum(mats):
    out = 0
    for i, m in enumerate(mats):
        term = np.eye(1)
        for j, mm in enumerate(mats):
            term = np.kron(term, m if i == j else np.eye(mm.shape[0]))
        out = out + term
    return out


def test_fdm_tables_and_interior_stencil(golden):
    # the hand tables are the FDM Kronecker sums (fdm.rs:141-150)
    lap1 = np.array([[2, -1], [-1, 2]])
    for d, table in golden["fdm_laplace_interior"]["cases"].items():
        assert (kron_sum([lap1] * int(d)) == np.array(table)).all()
    # FEEC interior = FDM stencil (fdm.rs:166-225), grid [0, n]^d with h = 1
    for d in range(1, 5):
        for nb in range(1, 3 if d == 4 else 4):
            n = nb + 2
            cells, coords = o.kuhn(d, n)
            coords = coords * n
            cx = o.Complex(d, cells)
            L = cx.lengths_from_coords(coords)
            A = o.assemble(cx, L, o.DIF_BOTH, 1).todense()
            M = o.assemble(cx, L, o.MASS, 0).todense()
            b = M @ np.ones(M.shape[0])
            A = A / b[:, None]
            inner = ~np.any((coords == 0) | (coords == n), axis=1)
            got = A[np.ix_(inner, inner)]
            assert np.abs(got - np.round(got)).max() < 1e-10
            one = np.diag(np.full(n - 1, 2)) - np.diag(np.ones(n - 2), 1) - np.diag(np.ones(n - 2), -1)
            assert (np.round(got) == kron_sum([one] * d)).all(), (d, nb)


# ---------------------------------------------------------------------------
# closed forms that do not go through the oracle's formula
# ---------------------------------------------------------------------------

def random_simplex(d, rng):
    while True:
        x = rng.standard_normal((d + 1, d))
        if abs(np.linalg.det(x[1:] - x[0])) > 0.3:
            return x


def lengths_of(x):
    d = x.shape[0] - 1
    return np.array([np.sum((x[j] - x[i]) ** 2) for i, j in o.local_subsets(d, 1)])


def bary_grads(x):
    d = x.shape[0] - 1
    E = (x[1:] - x[0]).T  # columns = edge vectors
    Einv = np.linalg.inv(E)
    g = np.zeros((d + 1, d))
    g[1:] = Einv
    g[0] = -Einv.sum(axis=0)
    return g


def test_edge_element_mass_vector_calculus():
    """k=1 Whitney mass from W_ij = l_i grad l_j - l_j grad l_i and the exact
    barycentric moments: an independent textbook route (Nedelec edge mass)."""
    rng = np.random.default_rng(0)
    for d in (2, 3, 4):
        x = random_simplex(d, rng)
        vol = abs(np.linalg.det(x[1:] - x[0])) / math.factorial(d)
        g = bary_grads(x)
        n = d + 1
        Q = (np.ones((n, n)) + np.eye(n)) / (n * (n + 1)) * vol
        edges = o.local_subsets(d, 1)
        M = np.zeros((len(edges), len(edges)))
        for a, (i, j) in enumerate(edges):
            for b, (k, l) in enumerate(edges):
                M[a, b] = (Q[i, k] * g[j] @ g[l] - Q[i, l] * g[j] @ g[k]
                           - Q[j, k] * g[i] @ g[l] + Q[j, l] * g[i] @ g[k])
        np.testing.assert_allclose(o.element_matrix(d, o.MASS, 1, lengths_of(x)), M, rtol=1e-12, atol=1e-14)


def test_dif_both_is_gramian_of_curls():
    """dif_both(2) in 3D = Gram of curl W_ij = 2 grad l_i x grad l_j (operators.rs:977-997)."""
    rng = np.random.default_rng(1)
    x = random_simplex(3, rng)
    vol = abs(np.linalg.det(x[1:] - x[0])) / 6
    g = bary_grads(x)
    edges = o.local_subsets(3, 1)
    curls = [2 * np.cross(g[i], g[j]) for i, j in edges]
    G = np.array([[ci @ cj for cj in curls] for ci in curls]) * vol
    np.testing.assert_allclose(o.element_matrix(3, o.DIF_BOTH, 2, lengths_of(x)), G, rtol=1e-12, atol=1e-14)


def test_top_mass_is_inverse_volume():
    rng = np.random.default_rng(2)
    for d in range(1, 5):
        x = random_simplex(d, rng)
        vol = abs(np.linalg.det(x[1:] - x[0])) / math.factorial(d)
        np.testing.assert_allclose(o.element_matrix(d, o.MASS, d, lengths_of(x)), [[1.0 / vol]], rtol=1e-12)


def test_metric_roundtrip_and_cayley_menger():
    """SimplexLengthsSq::from_metric . metric = id (lengths/simplex.rs:351-373)
    and CM = det g / (d!)^2 up to sign (lengths/simplex.rs:182-188)."""
    rng = np.random.default_rng(3)
    for d in range(1, 5):
        for sig in ("riem", "mink"):
            A = rng.standard_normal((d, d))
            g = A @ A.T + d * np.eye(d)
            if sig == "mink":
                g[0, :] *= 1
                g = A.T @ np.diag([-1.0] + [1.0] * (d - 1)) @ A
            edges = o.local_subsets(d, 1)
            s = []
            for i, j in edges:
                if i == 0:
                    s.append(g[j - 1, j - 1])
                else:
                    s.append(g[i - 1, i - 1] + g[j - 1, j - 1] - 2 * g[i - 1, j - 1])
            gg, gi, det = o.cell_metric(d, np.array(s))
            np.testing.assert_allclose(gg, g, rtol=1e-12, atol=1e-12)
            np.testing.assert_allclose(gi @ gg, np.eye(d), atol=1e-10)
            D = np.zeros((d + 2, d + 2))
            for (i, j), v in zip(edges, s):
                D[i, j] = D[j, i] = v
            D[d + 1, : d + 1] = D[: d + 1, d + 1] = 1.0
            cm = (-1) ** (d + 1) / math.factorial(d) ** 2 / 2 ** d * np.linalg.det(D)
            np.testing.assert_allclose(abs(cm), abs(det) / math.factorial(d) ** 2, rtol=1e-9)


# ---------------------------------------------------------------------------
# topology laws
# ---------------------------------------------------------------------------

def stirling2(n, k):
    return sum((-1) ** i * math.comb(k, i) * (k - i) ** n for i in range(k + 1)) // math.factorial(k)


@pytest.mark.parametrize("d,n", [(1, 5), (2, 4), (3, 3), (4, 2)])
def test_skeleton_counts_colex_and_euler(d, n):
    cells, _ = o.kuhn(d, n)
    perm = np.random.default_rng(d).permutation(len(cells))
    cx = o.Complex(d, cells[perm])
    chi = 0
    for k in range(d + 1):
        closed = sum(math.comb(d, m) * math.factorial(k) * stirling2(m, k) * n ** m * (n + 1) ** (d - m)
                     for m in range(k, d + 1))
        assert cx.nsimplices(k) == closed
        sk = cx.skeleton(k)
        brute = sorted({tuple(c[list(s)]) for c in np.sort(cells, axis=1) for s in itertools.combinations(range(d + 1), k + 1)},
                       key=lambda t: t[::-1])
        assert [tuple(r) for r in sk] == brute
        chi += (-1) ** k * cx.nsimplices(k)
    assert chi == 1
    # local faces map to the simplices they are
    cl = cx.cells()
    for k in range(d + 1):
        sk = cx.skeleton(k)
        cf = cx.cell_faces(k)
        loc = o.local_subsets(d, k)
        for c in range(0, len(cl), max(1, len(cl) // 50)):
            for f, s in enumerate(loc):
                assert tuple(sk[cf[c, f]]) == tuple(cl[c][s])


@pytest.mark.parametrize("d", [1, 2, 3, 4])
def test_dd_zero_and_boundary(d):
    cells, coords = o.kuhn(d, 2)
    cx = o.Complex(d, cells)
    for k in range(d - 1):
        dd = (cx.coboundary(k + 1).to_scipy() @ cx.coboundary(k).to_scipy()).toarray()
        assert (dd == 0).all()
    # boundary k-simplices: those whose vertices all lie on one face of the box
    for k in range(d + 1):
        flags = cx.boundary_flags(k)
        sk = cx.skeleton(k)
        X = coords[sk]  # [N, k+1, d]
        on = np.zeros(len(sk), bool)
        for a in range(d):
            on |= np.all(X[:, :, a] == 0, axis=1) | np.all(X[:, :, a] == 1, axis=1)
        assert (flags.astype(bool) == on).all()
    o_signs = cx.orientation()
    assert o_signs is not None and set(np.unique(o_signs)) <= {-1, 1}


# ---------------------------------------------------------------------------
# assembly route laws (whitney_complex.rs:734-861)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("d", [2, 3])
def test_route_laws(d):
    cells, coords = jittered(d, 3, 0.15, 7)
    cx = o.Complex(d, cells)
    L = cx.lengths_from_coords(coords)
    for rel in (False, True):
        c = o.WhitneyComplex(cx, L, relative=rel)
        for k in range(d):
            D = c.dif(k).to_scipy()
            M = c.mass(k + 1).to_scipy()
            lhs = c.dif_both(k + 1).to_scipy().toarray()
            np.testing.assert_allclose(lhs, (D.T @ M @ D).toarray(), atol=1e-12 * abs(lhs).max())
            lhs = c.dif_test(k + 1).to_scipy().toarray()
            np.testing.assert_allclose(lhs, (D.T @ M).toarray(), atol=1e-12 * abs(lhs).max())
            lhs = c.dif_trial(k + 1).to_scipy().toarray()
            np.testing.assert_allclose(lhs, (M @ D).toarray(), atol=1e-12 * abs(lhs).max())
    full = o.WhitneyComplex(cx, L)
    rel = o.WhitneyComplex(cx, L, relative=True)
    for k in range(d + 1):
        keep = rel.interior(k).astype(bool)
        a = full.mass(k).to_scipy().toarray()[np.ix_(keep, keep)]
        np.testing.assert_array_equal(rel.mass(k).to_scipy().toarray(), a)


def test_exact_zero_entries_are_dropped():
    """galerkin.rs:173: entries that are exactly 0.0 in every cell are never
    pushed; on an unjittered 2D grid the k=1 mass has such entries."""
    cells, coords = o.kuhn(2, 2)
    cx = o.Complex(2, cells)
    L = cx.lengths_from_coords(coords)
    m = o.assemble(cx, L, o.MASS, 1)
    el = o.element_matrices(cx, L, o.MASS, 1)
    assert (el == 0).any()
    structural = set()
    faces = cx.cell_faces(1)
    for c in range(len(el)):
        for i in range(3):
            for j in range(3):
                if el[c, i, j] != 0.0:
                    structural.add((faces[c, i], faces[c, j]))
    rows = np.repeat(np.arange(m.shape[0]), np.diff(m.row_ptr))
    assert set(zip(rows.tolist(), m.col.tolist())) == structural


# ---------------------------------------------------------------------------
# solves (elliptic.rs:260-287, fe.rs:191-215)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("d", [1, 2, 3])
def test_mixed_solve_matches_direct(d):
    cells, coords = o.kuhn(d, 2)
    cx = o.Complex(d, cells)
    L = cx.lengths_from_coords(coords)
    for rel in (True, False):
        c = o.WhitneyComplex(cx, L, relative=rel)
        for k in range(d + 1):
            f = np.array([(i % 7) - 3.0 for i in range(cx.nsimplices(k))])
            h = c.harmonics(k)
            kkt, rhs, ns, nu = o.assemble_mixed_kkt(c, f, k, h)
            x_lu = spla.spsolve(kkt.to_scipy().tocsc(), rhs)
            for pre in ("jacobi", "direct"):
                sig, u, p, rep = o.solve_source(c, f, k, pre, rtol=1e-11, max_iters=100000)
                assert rep.converged
                x = np.concatenate([sig[c.inclusion_indices(k - 1)] if k > 0 else [], u[c.inclusion_indices(k)], p])
                assert np.linalg.norm(x - x_lu) / max(np.linalg.norm(x_lu), 1.0) < 1e-7, (rel, k, pre)


def test_cg_matches_cholesky():
    cells, coords = jittered(2, 6, 0.2, 1)
    cx = o.Complex(2, cells)
    L = cx.lengths_from_coords(coords)
    M = o.assemble(cx, L, o.MASS, 1)
    b = np.sin(np.arange(M.shape[0]) * 0.5 + 0.5)
    x, rep = o.cg(M, b, 1.0 / M.diagonal(), 1e-12)
    assert rep.converged and rep.iters < 60
    np.testing.assert_allclose(x, spla.spsolve(M.to_scipy().tocsc(), b), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("m", [2, 5, 8, 16])
def test_chebyshev_recurrence_is_the_chebyshev_residual_polynomial(m):
    """1 - lambda p(lambda) = T_m((theta-lambda)/delta) / T_m(theta/delta) on a
    diagonal operator, with T_m from numpy's Chebyshev series (not the
    recurrence): a wrong coefficient or a dropped direction breaks it."""
    lam = np.concatenate([np.linspace(1e-3, 1.0, 40), [0.37, 0.999]])
    c0, steps = o.chebyshev_coefficients(1.0, m)
    b = o.CHEB_LMAX_MARGIN
    a = b * o.cheb_interval_fraction(m)
    theta, delta = (b + a) / 2, (b - a) / 2
    P = sp.diags(lam).tocsr()
    z = o.chebyshev_apply(P, np.ones_like(lam), np.ones_like(lam), c0, steps)
    Tm = np.polynomial.chebyshev.chebval
    coef = [0] * m + [1]
    want = Tm((theta - lam) / delta, coef) / Tm(theta / delta, coef)
    np.testing.assert_allclose(1 - lam * z, want, rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("m", [1, 2, 8, 32])
def test_chebyshev_three_term_form_matches_two_term(m):
    """The three-term recurrence on D^-1/2 P D^-1/2 with the unit diagonal
    implied (the mixed GPU path's form) gives the two-term recurrence's
    z = p(D^-1 P) D^-1 r on a real hdif_gram(1) block."""
    cells, coords = jittered(3, 3, 0.1, 4)
    cx = o.Complex(3, cells)
    c = o.WhitneyComplex(cx, cx.lengths_from_coords(coords))
    P = c.hdif_gram(1).to_scipy().tocsr()
    dinv = 1.0 / P.diagonal()
    lmax = np.abs(np.linalg.eigvals((dinv[:, None] * P.toarray()))).max()
    c0, steps = o.chebyshev_coefficients(lmax, m)
    r = np.cos(np.arange(P.shape[0]) * 1.3)
    z2 = o.chebyshev_apply(P, dinv, r, c0, steps)
    z3 = o.chebyshev_apply_three_term(P, dinv, r, c0, steps)
    np.testing.assert_allclose(z3, z2, rtol=1e-9, atol=1e-12 * np.abs(z2).max())


def test_chebyshev_jacobi_preconditioner():
    """lambda_max of D^-1 hdif_gram per block against a dense eigensolve, and
    a preconditioned MINRES that needs far fewer iterations than Jacobi."""
    cells, coords = jittered(2, 12, 0.1, 0)
    cx = o.Complex(2, cells)
    L = cx.lengths_from_coords(coords)
    c = o.WhitneyComplex(cx, L)
    prec = o.chebyshev_jacobi(c, 1, 0, 8)
    for b, g in enumerate((c.hdif_gram(0), c.hdif_gram(1))):
        A = g.todense()
        dh = 1 / np.sqrt(np.diag(A))
        lmax = np.linalg.eigvalsh(dh[:, None] * A * dh[None, :]).max()
        assert abs(prec.lambda_max[b] - lmax) <= 0.02 * lmax, (b, prec.lambda_max[b], lmax)
    f = np.sin(np.arange(cx.nsimplices(1)) * 0.7)
    kkt, rhs, ns, nu = o.assemble_mixed_kkt(c, f, 1, c.harmonics(1))
    A = kkt.to_scipy()
    x, rep = o.minres_py(lambda v: A @ v, prec, rhs, 1e-10, 20000)
    xj, repj = o.minres(kkt, rhs, o.mixed_jacobi(c, 1, 0), 1e-10, 20000)
    assert rep.converged and repj.converged
    assert rep.iters * 4 < repj.iters, (rep.iters, repj.iters)
    x_lu = spla.spsolve(A.tocsc(), rhs)
    assert np.linalg.norm(x - x_lu) / np.linalg.norm(x_lu) < 1e-7
    # the C restatement of the same preconditioned MINRES
    xc, repc = o.minres_cheb(kkt, rhs, prec, 1e-10, 20000)
    assert repc.converged and abs(repc.iters - rep.iters) <= 2, (repc.iters, rep.iters)
    assert np.linalg.norm(xc - x) / np.linalg.norm(x) < 1e-8


def test_minres_finite_termination():
    """krylov.rs:225-319: symmetric indefinite from a known spectrum."""
    rng = np.random.default_rng(5)
    n = 24
    Qm, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Qm @ np.diag(np.linspace(-3, 4, n) + 0.05) @ Qm.T
    A = (A + A.T) / 2
    Acsr = o.Csr((n, n), np.arange(n + 1) * n, np.tile(np.arange(n), n), A.reshape(-1))
    b = rng.standard_normal(n)
    x, rep = o.minres(Acsr, b, None, 1e-12, 2 * n)
    assert rep.converged
    np.testing.assert_allclose(A @ x, b, atol=1e-9)


# ---------------------------------------------------------------------------
# Hodge-Dirac (dirac.rs:662-676, 791-815)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("d", [2, 3])
def test_dirac_symmetry_classes(d):
    cells, coords = jittered(d, 2, 0.1, 3)
    cx = o.Complex(d, cells)
    L = cx.lengths_from_coords(coords)
    c = o.WhitneyComplex(cx, L)
    _, mb, skew = o.hodge_dirac(c, -1.0)
    A = skew.to_scipy()
    assert abs(A + A.T).max() == 0.0
    _, _, sa = o.hodge_dirac(c, 1.0)
    B = sa.to_scipy()
    assert abs(B - B.T).max() == 0.0
    Mb = mb.to_scipy()
    assert abs(Mb - Mb.T).max() < 1e-14


# ---------------------------------------------------------------------------
# the checks have teeth: each planted mistake breaks a known answer
# ---------------------------------------------------------------------------

MUTATIONS = {1: "barycentric Gramian without the diagonal weight 2",
             2: "deletion sign dropped in the Whitney expansion",
             3: "wrong sign of s_ij in the polarization",
             4: "local boundary operator with flipped signs",
             5: "cell volume without 1/d!"}


@pytest.mark.parametrize("m", sorted(MUTATIONS))
def test_known_answers_catch_planted_mistakes(m, golden, tmp_path):
    so = tmp_path / f"liboracle_m{m}.so"
    subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-shared", f"-DORC_MUTATION={m}", "-o", str(so),
                    os.path.join(ROOT, "oracle", "feec_oracle.c"), "-lm"], check=True)
    L = o.lib(str(so))
    failures = 0
    for d, expected in golden["laplacian_refcell"]["cases"].items():
        got = o.element_matrix(int(d), o.DIF_BOTH, 1, refcell(int(d)), L=L)
        failures += not np.allclose(got, expected, rtol=1e-12)
    for d, expected in golden["mass_refcell"]["cases"].items():
        got = o.element_matrix(int(d), o.MASS, 0, refcell(int(d)), L=L)
        failures += not np.allclose(got, expected, rtol=1e-12)
    s = refcell(2)
    for kind, key in ((o.MASS, "hodge_mass_dim2_grade1"), (o.DIF_TRIAL, "dif_trial_n2_k1"),
                      (o.DIF_TEST, "dif_test_n2_k1")):
        failures += not np.allclose(o.element_matrix(2, kind, 1, s, L=L), golden[key]["value"], rtol=1e-12)
    assert failures > 0, MUTATIONS[m]


def test_chebyshev_interval_top_bounds_lambda_max():
    """The interval's top is never below lambda_max(D^-1 hdif_gram) (dense
    eigensolve), also on the solid-torus slab whose vertex block 30 power
    steps underestimate by 6% (1.05 times that estimate would not cover it),
    and the Chebyshev-preconditioned MINRES there needs more than one step
    (a polynomial that is not positive on the spectrum breaks the Lanczos
    recurrence at once)."""
    from _homology_meshes import FIXTURES, squared_lengths
    cases = []
    dim, cells, geom = FIXTURES["tunnel"]()
    cx = o.Complex(dim, cells)
    cases.append((cx, squared_lengths(cx, geom), dim))
    cells, coords = jittered(2, 8, 0.1, 1)
    cx2 = o.Complex(2, cells)
    cases.append((cx2, cx2.lengths_from_coords(coords), 2))
    for cx, L, dim in cases:
        c = o.WhitneyComplex(cx, L)
        for k in range(dim + 1):
            blocks = ([c.hdif_gram(k - 1)] if k else []) + [c.hdif_gram(k)]
            P = [g.todense() for g in blocks]
            for A in P:
                d = np.diag(A)
                lmax = np.linalg.eigvals(A / d[:, None]).real.max()
                rows = (np.abs(A).sum(1) / d).max()
                assert rows >= lmax * (1 - 1e-12)
            prec = o.chebyshev_jacobi(c, k, 0, 16)
            for b, A in enumerate(P):
                bi = b if k else 1
                d = np.diag(A)
                lmax = np.linalg.eigvals(A / d[:, None]).real.max()
                top = o.cheb_interval_top(prec.lambda_max[bi], (np.abs(A).sum(1) / d).max())
                assert top >= lmax * (1 - 1e-12), (k, b, top, lmax)
    c = o.WhitneyComplex(cases[0][0], cases[0][1])
    f = np.cos(1.0 + np.arange(cases[0][0].nsimplices(0)))
    kkt, rhs, ns, nu = o.assemble_mixed_kkt(c, f, 0, c.harmonics(0))
    x, rep = o.minres_cheb(kkt, rhs, o.chebyshev_jacobi(c, 0, 1, 16))
    assert rep.converged and rep.iters > 1
    r = kkt.to_scipy() @ x - rhs
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(rhs)
