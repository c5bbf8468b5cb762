"""feec_csr_cg_noalloc (the Jacobi CG on the library's fixed device scratch)
against feec_csr_cg, through the C ABI: the same iterates bit for bit, on two
calls in a row, the second with a smaller system, and no device allocation."""
import numpy as np
import pytest

import formoniq_amd as fa

pytestmark = pytest.mark.gpu


def _gram(ctx, d, n, k, seed):
    g = fa.CartesianGrid(d, n)
    cells, coords = g.triangulate()
    coords = g.jitter(coords, 0.05, seed)
    topo = fa.Complex(d, cells, nverts=len(coords), ctx=ctx)
    geo = fa.MeshLengthsSq.from_coords(topo, coords)
    return fa.WhitneyComplex(geo).hdif_gram(k)


def test_noalloc_cg_matches_cg_on_two_calls():
    ctx = fa.Context(0)
    big = _gram(ctx, 3, 16, 1, 1)   # 31,024 rows: H(curl) Gram of a jittered 16^3 Kuhn mesh
    small = _gram(ctx, 2, 8, 0, 2)  # 81 rows, after the big one: stale scratch must not leak
    assert big.shape[0] == 31024 and small.shape[0] == 81
    for A, seed in ((big, 3), (small, 4)):
        b = fa.normal_vector(A.shape[0], seed)
        before = ctx.allocations()
        x1, r1 = A.cg_noalloc(b, rtol=1e-10, max_iters=20000)
        assert ctx.allocations() == before  # nothing allocated
        x0, r0 = A.cg(b, rtol=1e-10, max_iters=20000)
        assert ctx.allocations() > before   # the allocating entry point does allocate
        assert r1.converged and r0.converged
        assert r1.iters == r0.iters and r1.residual == r0.residual
        np.testing.assert_array_equal(x1, x0)


def test_noalloc_cg_rejects_systems_beyond_its_scratch():
    ctx = fa.Context(0)
    A = _gram(ctx, 2, 400, 1, 5)  # 480,800 rows > 262,144
    with pytest.raises(fa.FeecError) as e:
        A.cg_noalloc(np.ones(A.shape[0]))
    assert e.value.name == "FEEC_E_ARG" and "262144" in str(e.value)
