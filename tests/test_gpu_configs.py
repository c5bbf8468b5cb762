"""The configurations at their workloads on the GPU.

C3: the 4D Lorentzian 16^4 Kuhn box (1,572,864 cells): every all-grade element
    block of every cell bit-identical to the oracle's, and the Hodge-Dirac
    assembly checked against the oracle's digests or, failing those, by its
    symmetry and structure.
C4: the 512-mesh batch of seed 4 (d in 2..4, k in 0..d-1, 1e3..1e5 cells) in
    one feec_batch_run: every request certified to 1e-10, and one request of
    each dimension bit-identical to the same request solved alone.
C5: a compressed serving stream that includes a ~1.2e6-cell request: that
    request's solution is bit-identical alone and in the mix."""
import json
import os
import time

import numpy as np
import pytest

import formoniq_amd as fa
from oracle import oracle as o

from bench import config_requests

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def c3():
    g = fa.CartesianGrid(4, 16)
    cells, coords = g.triangulate()
    coords = g.jitter(coords, 0.02, 3, axes=[1, 2, 3])
    gcx = fa.Complex(4, cells, nverts=len(coords))
    geo = fa.MeshLengthsSq.from_coords(gcx, coords, True)
    ocx = o.Complex(4, cells)
    oL = ocx.lengths_from_coords(coords, True)
    assert np.array_equal(geo.lengths(), oL)
    return gcx, geo, ocx, oL


C3_BLOCKS = [(fa.MASS, k) for k in range(5)] + [(fa.DIF_TRIAL, k) for k in range(1, 5)]


@pytest.mark.parametrize("kind,k", C3_BLOCKS)
def test_c3_element_matrices_bit_exact(c3, kind, k):
    gcx, geo, ocx, oL = c3
    assert gcx.nsimplices(4) == 1572864
    g = fa.WhitneyComplex(geo).element(kind, k)
    e = o.element_matrices(ocx, oL, kind, k)
    assert g.shape == e.shape
    bad = int((g != e).sum())
    assert bad == 0, (kind, k, bad)


def _csr_digests(a):
    import hashlib
    rp, col, val = a.host()
    return {"row_ptr": hashlib.sha256(np.ascontiguousarray(rp, "<i8").tobytes()).hexdigest(),
            "col": hashlib.sha256(np.ascontiguousarray(col, "<i4").tobytes()).hexdigest(),
            "val": hashlib.sha256(np.ascontiguousarray(val, "<f8").tobytes()).hexdigest()}


def test_c3_dirac_assembly(c3):
    """Self-adjoint Hodge-Dirac on the 16^4 Lorentzian box: offsets, counts
    within the structural counts (184,520,961 and 316,659,968), op = op^T
    exactly on every stored entry, and -- where the oracle's digests are
    committed -- mass_block and op bit-identical to the oracle's."""
    gcx, geo, ocx, oL = c3
    hd = fa.HodgeDirac.assemble_selfadjoint(fa.WhitneyComplex(geo))
    np.testing.assert_array_equal(hd.offsets, np.cumsum([0, 83521, 1102400, 3476480, 4030464, 1572864]))
    assert hd.mass_block.nnz <= 184520961 and hd.op.nnz <= 316659968
    rp, col, val = hd.op.host()
    rows = np.repeat(np.arange(len(rp) - 1, dtype=np.int64), np.diff(rp))
    key = rows * len(rp) + col
    tkey = col.astype(np.int64) * len(rp) + rows
    order = np.argsort(tkey, kind="stable")
    assert np.array_equal(tkey[order], key)   # the transposed pattern is the pattern
    assert np.array_equal(val[order], val)    # and the values match entry for entry
    path = os.path.join(HERE, "golden", "c3_dirac_digests.json")
    if os.path.exists(path):
        with open(path) as f:
            ref = json.load(f)
        assert hd.mass_block.nnz == ref["mass_block_nnz"] and hd.op.nnz == ref["op_nnz"]
        assert _csr_digests(hd.mass_block) == ref["mass_block"]
        assert _csr_digests(hd.op) == ref["op"]


def test_c4_batch_at_workload():
    reqs = config_requests(fa, 512, 4, [2, 3, 4], 1e3, 1e5)
    res = fa.batch_run(reqs)
    assert len(res) == 512
    bad = [(q.dim, q.grade, r.status, r.residual) for q, r in zip(reqs, res)
           if r.status != 0 or not r.residual <= 1e-10]
    assert not bad, bad[:5]
    sample = []
    for d in (2, 3, 4):
        sample.append(next(i for i, q in enumerate(reqs) if q.dim == d))
    for i in sample:
        alone = fa.batch_run([reqs[i]])[0]
        r = res[i]
        assert (r.iters, r.refinements, r.residual) == (alone.iters, alone.refinements, alone.residual), i
        assert r.u_norm == alone.u_norm and r.u_checksum == alone.u_checksum, i


def _serve(reqs, inflight, gaps=None):
    opts = fa.ServerOpts()
    opts.hbm_budget_bytes = 64 << 30
    opts.max_inflight = inflight
    opts.solve = fa.solve_opts()
    srv = fa.Server(opts)
    tickets = []
    for i, r in enumerate(reqs):
        if gaps is not None:
            time.sleep(gaps[i])
        tickets.append(srv.submit(r))
    srv.drain()
    out = [(srv.poll(t), srv.solution(t)) for t in tickets]
    srv.close()
    return out


def test_c5_stream_with_a_million_cell_request():
    big = fa.Request()
    big.dim, big.grade, big.n, big.bc = 3, 1, 58, fa.FULL  # 58^3 * 6 = 1,170,672 cells
    big.jitter = 0.05
    big.mesh_seed, big.source_seed = 500001, 550001
    big.precision = fa.FP64_PC32
    (r_alone, u_alone), = _serve([big], 1)
    assert r_alone.status == 0 and r_alone.residual <= 1e-10 and r_alone.ncells == 1170672
    stream = config_requests(fa, 12, 5, [2, 3], 1e3, 2e5)
    reqs = stream[:6] + [big] + stream[6:]
    gaps = np.random.default_rng(5).exponential(1.0 / 20.0, size=len(reqs))
    mixed = _serve(reqs, 4, gaps)
    r_mix, u_mix = mixed[6]
    assert r_mix.status == 0
    np.testing.assert_array_equal(u_mix, u_alone)
    for r, _ in mixed:
        assert r.status == 0 and r.residual <= 1e-10
    spans = sorted((r.admit_time, r.finish_time) for r, _ in mixed)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
