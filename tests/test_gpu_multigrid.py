"""The multigrid / auxiliary-space block preconditioner on the GPU against the
oracle (oracle.hx_block_preconditioner, the restatement of hx.rs and
multigrid.rs), and the reference's own law for it: the block-MINRES iteration
count grows at most slowly under refinement (elliptic.rs:338-367 states
fine <= coarse + 8 for its exact block solves, 2D k = 1 on the relative
complex of the unit square)."""
import time

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import formoniq_amd as fa
from oracle import oracle as o

from _common import both

pytestmark = pytest.mark.gpu


def mg_opts(**kw):
    return fa.solve_opts(preconditioner=fa.PC_MULTIGRID, **kw)


CASES = [(2, 12, 0.1, 1, False), (2, 13, 0.1, 1, True), (2, 10, 0.1, 0, False), (2, 9, 0.1, 2, True),
         (3, 6, 0.05, 1, False), (3, 5, 0.05, 1, True), (3, 5, 0.05, 2, False), (3, 6, 0.05, 0, True)]


@pytest.mark.parametrize("m", [1, 4])
@pytest.mark.parametrize("d,n,amp,k,rel", CASES)
def test_hx_application_matches_oracle(d, n, amp, k, rel, m):
    """z = B r with B = diag(HX(k-1), HX(k), I) (m = 1), or each block
    accelerated by the Chebyshev polynomial in B hdif_gram with m
    applications of B: FP64 equal to the oracle's to 1e-10 (summation order
    and the coarse inverse's method differ). The FP32 path (its Chebyshev
    steps multiply by the exactly symmetric FP32 copy of hdif_gram, which the
    oracle models with single_symmetric) to 1e-5 (m = 1) or 1e-4 (m = 4: four
    cycles and three products in FP32). The lambda_max estimates (6 Lanczos
    steps from the same start) equal the oracle's to 1e-9 on the FP64 path
    and to 1e-4 on the FP32 path."""
    ocx, oL, gcx, geo, cells, coords = both(d, n, amp, seed=n)
    gc = fa.WhitneyComplex(geo, relative=rel)
    prob = fa.Problem(gc, k, mg_opts(precision=fa.FP64, cheb_degree=m))
    prob.assemble()
    oc = o.WhitneyComplex(ocx, oL, relative=rel)
    ns, nu, nh = prob.sizes
    info = prob.preconditioner()
    assert info["degree"] == m
    lam = None
    if m > 1:
        lam = info["lambda_max"]
        want_lam = o.hx_block_lambdas(oc, coords, k, n)
        for b in range(2 if k > 0 else 1):
            i = b if k > 0 else 1
            assert abs(lam[i] - want_lam[i]) <= 1e-9 * want_lam[i], (lam, want_lam)
    prec = o.hx_block_preconditioner(oc, coords, k, n, nh, cheb_m=m, lam=lam)
    r = np.sin(np.arange(ns + nu + nh) * 0.7 + 0.3)
    want = prec(r)
    got = prob.apply_preconditioner(r, precision=fa.FP64)
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 1e-10 * scale, np.abs(got - want).max() / scale
    # the FP32 path: its own lambda_max (FP32 cycle) and the FP32 copy of P
    prob32 = fa.Problem(gc, k, mg_opts(precision=fa.FP64_PC32, cheb_degree=m))
    prob32.assemble()
    want32 = want
    if m > 1:
        lam32 = prob32.preconditioner()["lambda_max"]
        want_lam32 = o.hx_block_lambdas(oc, coords, k, n, single=True)
        for i in ([0, 1] if k > 0 else [1]):
            assert abs(lam32[i] - want_lam32[i]) <= 1e-4 * want_lam32[i], (lam32, want_lam32)
        want32 = o.hx_block_preconditioner(oc, coords, k, n, nh, cheb_m=m, lam=lam32, single=True)(r)
    tol32 = 1e-5 if m == 1 else 1e-4
    got32 = prob32.apply_preconditioner(r, precision=fa.MIXED)
    assert np.abs(got32 - want32).max() <= tol32 * scale, np.abs(got32 - want32).max() / scale
    got_pc32 = prob32.apply_preconditioner(r, precision=fa.FP64_PC32)
    assert np.abs(got_pc32 - want32).max() <= tol32 * scale


def _iters(d, n, k, rel, amp, precision=fa.FP64, m=1):
    ocx, oL, gcx, geo, _, _ = both(d, n, amp, seed=0)
    gc = fa.WhitneyComplex(geo, relative=rel)
    f = np.array([((i % 5) - 2.0) * 0.3 for i in range(gcx.nsimplices(k))])
    opts = mg_opts(precision=precision, max_refinements=0, cheb_degree=m)
    sig, u, p, rep = fa.solve_source(gc, f, k, opts, raise_on_failure=False)
    return rep


def test_block_minres_iteration_count_is_mesh_independent():
    """elliptic.rs:338-367 states its law (refinement 4 -> 8 adds at most 8
    MINRES iterations; 2D k = 1, relative complex of the unit square, rtol
    1e-10) for exact block solves, and the oracle's exact-block MINRES obeys
    it. The auxiliary-space blocks are spectrally equivalent to the exact ones
    with constants that settle only after a few refinements (the V-cycle's
    own contraction is h-independent, 0.23-0.31 from 16^2 to 64^2): from
    16^2 to 128^2 each doubling adds at most 16 iterations and the three
    doublings together at most 40% (an h-dependent preconditioner doubles the
    count per doubling). The oracle's count at 16^2 agrees to 10%."""
    counts = [_iters(2, n, 1, True, 0.0).iters for n in (16, 32, 64, 128)]
    print("2D k=1 relative iterations by n = 16, 32, 64, 128:", counts, flush=True)
    for coarse, fine in zip(counts, counts[1:]):
        assert fine <= coarse + 16, counts
    assert counts[-1] <= 1.4 * counts[0], counts
    accel = [_iters(2, n, 1, True, 0.0, m=4).iters for n in (16, 32, 64, 128)]
    print("  with the degree-3 Chebyshev acceleration (m = 4):", accel, flush=True)
    for coarse, fine in zip(accel, accel[1:]):
        assert fine <= coarse + 8, accel
    assert max(accel) <= 0.4 * min(counts), (accel, counts)
    ocx, oL, gcx, geo, cells, coords = both(2, 16, 0.0, seed=0)
    oc = o.WhitneyComplex(ocx, oL, relative=True)
    f = np.array([((i % 5) - 2.0) * 0.3 for i in range(ocx.nsimplices(1))])
    kkt, rhs, ns, nu = o.assemble_mixed_kkt(oc, f, 1, oc.harmonics(1))
    A = kkt.to_scipy()
    _, rep = o.minres_py(lambda v: A @ v, o.hx_block_preconditioner(oc, coords, 1, 16, 0), rhs, 1e-10, 5000)
    assert abs(counts[0] - rep.iters) <= max(3, rep.iters // 10), (counts[0], rep.iters)


def test_3d_iteration_count_is_bounded():
    """3D k = 1 on the jittered full complex: 8^3, 16^3, 32^3, 64^3, plain
    HX blocks and with the Chebyshev acceleration (m = 4)."""
    counts = [_iters(3, n, 1, False, 0.05).iters for n in (8, 16, 32, 64)]
    accel = [_iters(3, n, 1, False, 0.05, m=4).iters for n in (8, 16, 32, 64)]
    print("3D k=1 iterations by n = 8, 16, 32, 64:", counts, "with m = 4:", accel, flush=True)
    for coarse, fine in zip(counts, counts[1:]):
        assert fine <= coarse + 16, counts
    assert counts[-1] <= 1.4 * counts[0], counts
    for coarse, fine in zip(accel, accel[1:]):
        assert fine <= coarse + 8, accel


@pytest.mark.parametrize("precision", [fa.FP64, fa.MIXED, fa.FP64_PC32])
@pytest.mark.parametrize("d,n,k,rel", [(2, 24, 1, False), (3, 8, 1, False), (3, 7, 1, True),
                                       (3, 6, 2, False), (2, 20, 0, False)])
def test_multigrid_solve_is_certified(precision, d, n, k, rel):
    """solve_source with the multigrid preconditioner in every precision mode:
    the true residual of the oracle's FP64 KKT system <= 1e-10 after the
    solve's own certificate, and the solution equal to SuperLU's to 1e-7."""
    ocx, oL, gcx, geo, _, _ = both(d, n, 0.1 if d == 2 else 0.05, seed=3)
    gc = fa.WhitneyComplex(geo, relative=rel)
    oc = o.WhitneyComplex(ocx, oL, relative=rel)
    f = fa.normal_vector(ocx.nsimplices(k), 9)
    sig, u, p, rep = fa.solve_source(gc, f, k, mg_opts(precision=precision))
    assert rep.converged and rep.residual <= 1e-10, rep.as_dict()
    kkt, rhs, ns, nu = o.assemble_mixed_kkt(oc, f, k, oc.harmonics(k))
    x = np.concatenate([sig[oc.inclusion_indices(k - 1)] if k > 0 else [], u[oc.inclusion_indices(k)], p])
    if p.size:
        gh, oh = gc.harmonics(k), oc.harmonics(k)
        x[-p.size:] *= np.sign(gh[:, 0] @ oh[:, 0])
    A = kkt.to_scipy()
    assert np.linalg.norm(rhs - A @ x) / np.linalg.norm(rhs) <= 1e-9
    x_lu = spla.spsolve(A.tocsc(), rhs)
    assert np.linalg.norm(x - x_lu) / np.linalg.norm(x_lu) <= 1e-7


def test_multigrid_needs_a_kuhn_grid_and_coordinates():
    """A mesh whose vertices are renumbered is no longer recognised as a Kuhn
    grid: PC_MULTIGRID is FEEC_E_UNSUPPORTED, PC_AUTO falls back to the
    polynomial. A geometry given by lengths alone has no frame for the
    vector-nodal space."""
    d, n = 2, 40
    cells, coords = fa.CartesianGrid(d, n).triangulate()
    coords = fa.CartesianGrid(d, n).jitter(coords, 0.1, 1)
    perm = np.random.default_rng(0).permutation(len(coords))
    inv = np.argsort(perm)
    cells2, coords2 = inv[cells], coords[perm]
    gcx = fa.Complex(d, cells2, nverts=len(coords2))
    geo = fa.MeshLengthsSq.from_coords(gcx, coords2)
    f = fa.normal_vector(gcx.nsimplices(1), 2)
    with pytest.raises(fa.FeecError) as e:
        fa.solve_source(fa.WhitneyComplex(geo), f, 1, mg_opts())
    assert e.value.name == "FEEC_E_UNSUPPORTED"
    *_, rep = fa.solve_source(fa.WhitneyComplex(geo), f, 1, fa.solve_opts())
    assert rep.converged and rep.residual <= 1e-10
    gcx = fa.Complex(d, cells, nverts=len(coords))
    geo_c = fa.MeshLengthsSq.from_coords(gcx, coords)
    geo_l = fa.MeshLengthsSq.from_lengths(gcx, geo_c.lengths())
    with pytest.raises(fa.FeecError) as e:
        fa.solve_source(fa.WhitneyComplex(geo_l), f, 1, mg_opts())
    assert e.value.name == "FEEC_E_UNSUPPORTED"


def test_c2_full_size_multigrid():
    """The 64^3 k = 1 system: FP64_PC32 and FP64 with the multigrid
    preconditioner certified to 1e-10 in at most 30 MINRES iterations, and
    the same solution as the polynomial's (to the solve tolerance)."""
    _, _, gcx, geo, _, _ = both(3, 64, 0.05, seed=2)
    wc = fa.WhitneyComplex(geo)
    f = fa.normal_vector(gcx.nsimplices(1), 2)
    t = [time.perf_counter()]
    _, u_mg, _, rep_mg = fa.solve_source(wc, f, 1, mg_opts())
    t.append(time.perf_counter())
    _, u64, _, rep64 = fa.solve_source(wc, f, 1, mg_opts(precision=fa.FP64))
    t.append(time.perf_counter())
    _, u_ch, _, rep_ch = fa.solve_source(wc, f, 1, fa.solve_opts(preconditioner=fa.PC_CHEBYSHEV))
    t.append(time.perf_counter())
    print("64^3 seconds (setup + solve): multigrid FP64_PC32 %.3f FP64 %.3f chebyshev %.3f"
          % tuple(np.diff(t)), flush=True)
    for rep in (rep_mg, rep64, rep_ch):
        assert rep.converged and rep.residual <= 1e-10, rep.as_dict()
    # the mesh-independent count (15-17 with the default degree), far below the polynomial's
    assert rep_mg.iters <= 30 and rep64.iters <= 30, (rep_mg.iters, rep64.iters)
    assert np.linalg.norm(u_mg - u_ch) <= 1e-8 * np.linalg.norm(u_ch)
    assert np.linalg.norm(u64 - u_ch) <= 1e-8 * np.linalg.norm(u_ch)
    print("64^3 iterations: multigrid FP64_PC32", rep_mg.iters, "FP64", rep64.iters,
          "chebyshev", rep_ch.iters)
