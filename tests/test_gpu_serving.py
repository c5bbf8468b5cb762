"""Batched solves and the continuous-batching server on the GPU.

Each result is checked against a standalone solve_source of the same request,
and the server test shows a request's solution is bit-identical alone and when
it shares the device with other requests (every reduction has a fixed order
and no state is shared between requests)."""
import time

import numpy as np
import pytest

import formoniq_amd as fa

pytestmark = pytest.mark.gpu


def req(dim, grade, n, seed, precision=fa.MIXED, jitter=0.05):
    r = fa.Request()
    r.dim, r.grade, r.n, r.bc = dim, grade, n, fa.FULL
    r.jitter = jitter
    r.mesh_seed, r.source_seed = seed, seed + 1000
    r.precision = precision
    return r


def standalone_u(r):
    g = fa.CartesianGrid(r.dim, r.n)
    cells, coords = g.triangulate()
    coords = g.jitter(coords, r.jitter, r.mesh_seed)
    topo = fa.Complex(r.dim, cells, nverts=len(coords))
    geo = fa.MeshLengthsSq.from_coords(topo, coords)
    f = fa.normal_vector(topo.nsimplices(r.grade), r.source_seed)
    _, u, _, rep = fa.solve_source(fa.WhitneyComplex(geo), f, r.grade, fa.solve_opts(precision=fa.FP64))
    return u


def continued_req():
    # 2D 64^2 k=1, jitter 0.1 (mesh seed 94, source seed 31): an FP64_PC32 solve
    # whose first run falls short of the true residual and continues
    r = req(2, 1, 64, 94, fa.FP64_PC32, jitter=0.1)
    r.source_seed = 31
    return r


MIX = [req(2, 0, 12, 1), req(2, 1, 10, 2), req(3, 0, 5, 3), req(3, 1, 4, 4), req(3, 2, 4, 5),
       req(4, 0, 2, 6), req(4, 1, 2, 7), req(4, 3, 2, 8), req(2, 1, 16, 9, fa.FP64),
       req(3, 1, 5, 10, fa.FP64_PC32), req(2, 2, 12, 11, fa.FP64_PC32),
       req(2, 1, 96, 12, fa.FP64_PC32, jitter=0.1), continued_req()]


def checksum(u):
    return float(np.sum(u * ((np.arange(u.size) % 7) - 3)))


def test_batch_run_matches_standalone():
    res = fa.batch_run(MIX)
    assert len(res) == len(MIX)
    for r, q in zip(res, MIX):
        assert r.status == 0 and r.residual <= 1e-10, (q.dim, q.grade, r.as_dict())
        u = standalone_u(q)
        assert abs(r.u_norm - np.linalg.norm(u)) <= 1e-7 * np.linalg.norm(u)
        assert abs(r.u_checksum - checksum(u)) <= 1e-7 * np.abs(u).sum()


def test_batch_lockstep_is_bit_identical_to_single_solves():
    """feec_batch_run advances the MINRES runs of all its requests in lockstep
    (one launch per phase); every request still gets the blocks, partial sums
    and iterates of a solve on its own, so iteration counts, residuals and
    the solution's norm and checksum equal the server's single solves bit for
    bit (mixed and FP64, harmonic border rows included; FP64_PC32 requests run
    one after another beside the lockstep runs)."""
    res = fa.batch_run(MIX)
    single = run_server(MIX, 1)
    for r, (s_, _), q in zip(res, single, MIX):
        assert r.status == 0 and s_.status == 0
        assert (r.iters, r.refinements) == (s_.iters, s_.refinements), (q.dim, q.grade)
        assert r.residual == s_.residual, (q.dim, q.grade)
        assert r.u_norm == s_.u_norm and r.u_checksum == s_.u_checksum, (q.dim, q.grade)
    # the last request's run continued (lockstep in the batch, by itself in the server)
    assert res[-1].refinements == 1, res[-1].as_dict()


def run_server(reqs, inflight, budget=32 << 30):
    # the polynomial preconditioner, as the batch runs it (single solves equal lockstep ones)
    opts = fa.ServerOpts()
    opts.hbm_budget_bytes = budget
    opts.max_inflight = inflight
    opts.solve = fa.solve_opts(preconditioner=fa.PC_CHEBYSHEV)
    srv = fa.Server(opts)
    tickets = [srv.submit(r) for r in reqs]
    srv.drain()
    out = []
    for t in tickets:
        r = srv.poll(t)
        assert r is not None
        out.append((r, srv.solution(t)))
    srv.close()
    return out


def test_server_same_alone_and_in_the_mix():
    target = req(3, 1, 6, 42)
    (alone, u_alone), = run_server([target], 1)
    assert alone.status == 0 and alone.residual <= 1e-10
    others = [req(2, 1, 20, 50 + i) for i in range(3)] + [req(3, 2, 5, 60), req(4, 1, 2, 61)]
    mixed = run_server(others[:3] + [target] + others[3:], 4)
    r_mix, u_mix = mixed[3]
    assert r_mix.status == 0
    np.testing.assert_array_equal(u_mix, u_alone)  # bit-identical
    for r, _ in mixed:
        assert r.status == 0 and r.residual <= 1e-10
        assert r.finish_time >= r.admit_time >= 0 and r.latency > 0
    # requests did overlap: some admitted before another finished
    spans = sorted((r.admit_time, r.finish_time) for r, _ in mixed)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))


def test_server_admission_budget():
    opts = fa.ServerOpts()
    opts.hbm_budget_bytes = 256 << 20
    opts.max_inflight = 2
    opts.solve = fa.solve_opts()
    srv = fa.Server(opts)
    with pytest.raises(fa.FeecError) as e:
        srv.submit(req(3, 1, 64, 1))
    assert e.value.name == "FEEC_E_ADMISSION"
    t = srv.submit(req(2, 1, 8, 1))
    srv.drain()
    assert srv.poll(t).status == 0
    srv.close()


def test_server_release_bounds_held_state():
    """A stream of small requests whose results are read and released holds
    nothing afterwards: finished entries (result and host copy of u) are
    erased, so a long continuous-batching stream does not grow host memory."""
    opts = fa.ServerOpts()
    opts.max_inflight = 2
    opts.solve = fa.solve_opts()
    srv = fa.Server(opts)
    peak = 0
    for wave in range(3):
        tickets = [srv.submit(req(2, 1, 6, 100 + 4 * wave + i)) for i in range(4)]
        srv.drain()
        peak = max(peak, srv.held()[0])
        for t in tickets:
            assert srv.poll(t).status == 0
            assert srv.solution(t).size > 0
            srv.release(t)
            with pytest.raises(fa.FeecError) as e:
                srv.poll(t)
            assert e.value.name == "FEEC_E_ARG"
        assert srv.held() == (0, 0)
    assert peak == 4
    srv.close()
