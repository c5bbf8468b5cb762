"""The oracle's Hodge-Dirac time steppers, pinned by what the mathematics fixes.

Leapfrog (time.rs:292-420; the loop of solve_dirac_leapfrog, dirac.rs:415-434):
the reference's own conservation law for the staggered invariant
(dirac.rs:765-784), time reversibility, second-order convergence to the exact
flow exp(t M^-1 A) y0, and the Stormer-Verlet closed form of a scalar
oscillator. A wrong kick coefficient, a swapped coloring or a dropped mass
solve fails one of them.

Gauss-Legendre (LinearIrk, time.rs:191-274; solve_dirac, dirac.rs:383-403):
the collocation and quadrature conditions that define the tableau, the
reference's energy conservation law (dirac.rs:737-754), and orders 2 and 4
against the exact flow."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import oracle as o


def unit_complex(d, n=2, relative=False):
    cells, coords = o.kuhn(d, n)
    cx = o.Complex(d, cells)
    L = cx.lengths_from_coords(coords)
    return o.WhitneyComplex(cx, L, relative=relative), L


def seed_field(n):
    return np.sin(0.7 * np.arange(n) + 0.3) + 0.1


@pytest.mark.parametrize("d", [1, 2, 3])
def test_staggered_energy_is_conserved(d):
    # dirac.rs:765-784: unit 2-grid, dt = 0.1 h_min, 200 steps, 1e-9 relative
    c, L = unit_complex(d)
    offsets, _, _ = o.hodge_dirac(c, -1.0)
    dt = 0.1 * np.sqrt(L.min())
    y, e = o.dirac_leapfrog(c, dt, 200, seed_field(int(offsets[-1])))
    assert e[0] > 0.0
    assert np.abs(e - e[0]).max() <= 1e-9 * e[0]


def test_coloring_and_split():
    c, _ = unit_complex(3)
    offsets, mass, op = o.hodge_dirac(c, -1.0)
    color = o.grade_parity_coloring(offsets)
    assert color.sum() == (offsets[2] - offsets[1]) + (offsets[4] - offsets[3])
    o.Leapfrog(mass, op, color, 0.01)  # asserts the block structure
    with pytest.raises(AssertionError):  # one color: op couples within it
        o.Leapfrog(mass, op, np.zeros_like(color), 0.01)


@pytest.mark.parametrize("d,rel", [(1, False), (2, False), (2, True)])
def test_reversible_and_second_order(d, rel):
    c, L = unit_complex(d, 3, relative=rel)
    offsets, mass, op = o.hodge_dirac(c, -1.0)
    n = int(offsets[-1])
    y0 = seed_field(n)
    dt = 0.05 * np.sqrt(L.min())
    y, _ = o.dirac_leapfrog(c, dt, 40, y0)
    back, _ = o.dirac_leapfrog(c, -dt, 40, y)
    np.testing.assert_allclose(back, y0, rtol=0, atol=1e-11 * np.abs(y0).max())
    # the exact flow of M y' = A y
    M = mass.to_scipy().toarray()
    A = op.to_scipy().toarray()
    T = 40 * dt
    exact = sla.expm(T * np.linalg.solve(M, A)) @ y0
    errs = []
    for k in (1, 2, 4):
        yk, _ = o.dirac_leapfrog(c, dt / k, 40 * k, y0)
        errs.append(np.linalg.norm(yk - exact))
    assert errs[0] > 0
    assert 3.2 <= errs[0] / errs[1] <= 4.8 and 3.2 <= errs[1] / errs[2] <= 4.8, errs


def test_scalar_oscillator_closed_form():
    # time.rs tests' oscillator y = (x, v), A = [[0, 1], [-w^2, 0]], M = I: one
    # kick-drift-kick step is Stormer-Verlet, whose n-step map has
    # cos(theta) = 1 - (w dt)^2 / 2 and x_n = cos(n theta) for (x0, v0) = (1, 0)
    w, dt, n = 1.7, 0.05, 300
    mass = o.Csr((2, 2), np.array([0, 1, 2]), np.array([0, 1]), np.array([1.0, 1.0]))
    op = o.Csr((2, 2), np.array([0, 1, 2]), np.array([1, 0]), np.array([1.0, -w * w]))
    lf = o.Leapfrog(mass, op, np.array([False, True]), dt)
    y = np.array([1.0, 0.0])
    for _ in range(n):
        y = lf.step(y)
    theta = np.arccos(1.0 - (w * dt) ** 2 / 2.0)
    assert abs(y[0] - np.cos(n * theta)) <= 1e-12


@pytest.mark.parametrize("s", [1, 2])
def test_gauss_legendre_tableau_collocation_and_quadrature(s):
    # Gauss collocation: A c^(k-1) = c^k / k (k = 1..s) and b . c^(k-1) = 1/k
    # (k = 1..2s), which fix A, b and c
    a, b, c = o.gauss_legendre(s)
    for k in range(1, s + 1):
        np.testing.assert_allclose(a @ c ** (k - 1), c ** k / k, rtol=0, atol=1e-15)
    for k in range(1, 2 * s + 1):
        assert abs(b @ c ** (k - 1) - 1.0 / k) <= 1e-15


@pytest.mark.parametrize("d", [1, 2, 3])
def test_gauss_legendre_conserves_energy(d):
    # dirac.rs:737-754: unit 2-grid, dt = 0.05, 100 steps, 1/2 y^T M y to 1e-9
    c, _ = unit_complex(d)
    offsets, _, _ = o.hodge_dirac(c, -1.0)
    y, e = o.dirac_irk(c, 0.05, 100, seed_field(int(offsets[-1])))
    assert e[0] > 0.0
    assert np.abs(e - e[0]).max() <= 1e-9 * e[0]


@pytest.mark.parametrize("s,order", [(1, 2), (2, 4)])
def test_gauss_legendre_order(s, order):
    c, L = unit_complex(2, 3)
    offsets, mass, op = o.hodge_dirac(c, -1.0)
    y0 = seed_field(int(offsets[-1]))
    M = mass.to_scipy().toarray()
    A = op.to_scipy().toarray()
    dt, n = 0.01, 8  # dt times the spectral radius of M^-1 A (18) well below 1
    exact = sla.expm(n * dt * np.linalg.solve(M, A)) @ y0
    errs = [np.linalg.norm(o.dirac_irk(c, dt / k, n * k, y0, stages=s)[0] - exact) for k in (1, 2, 4)]
    ratio = 2.0 ** order
    assert 0.75 * ratio <= errs[0] / errs[1] <= 1.25 * ratio, errs
    assert 0.75 * ratio <= errs[1] / errs[2] <= 1.25 * ratio, errs
