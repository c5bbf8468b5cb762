"""The autotuner's choices change no result bit.

Every candidate of a tuned launch choice is forced in turn (feec_tune_mode)
and the assembled matrices, the Hodge-Dirac operator and a batch of solves
are compared bit for bit; then the tuned mode fills the per-device table.
"""
import numpy as np
import pytest

import formoniq_amd as fa

from _common import both

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def restore_mode():
    yield
    fa.tune_mode(0)


def _csr(a):
    rp, col, val = a.host()
    return rp.copy(), col.copy(), val.copy()


def _same(x, y):
    for a, b in zip(x, y):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("d,n,k", [(2, 6, 1), (3, 3, 1), (3, 3, 2), (4, 2, 2)])
def test_element_variants_are_bit_identical(d, n, k):
    _, _, gcx, geo, _, _ = both(d, n, 0.1, seed=d)
    wc = fa.WhitneyComplex(geo)
    out = {}
    # mode 1: fused elements, per-lane segmented sums; 2: split by grade, the
    # wavefront's shared gathers in XCD runs (stage 512); 3: split, the same
    # with a stage of 1,024
    for mode in (1, 2, 3):
        fa.tune_mode(mode)
        kkt, _ = wc.mixed_kkt(k)
        hd = fa.HodgeDirac.assemble(wc)
        out[mode] = (_csr(kkt), _csr(hd.mass_block), _csr(hd.op))
    for m in (2, 3):
        for a, b in zip(out[1], out[m]):
            _same(a, b)


def test_batched_grid_candidates_are_bit_identical():
    def req(dim, grade, n, seed):
        r = fa.Request()
        r.dim, r.grade, r.n, r.bc = dim, grade, n, fa.FULL
        r.jitter, r.mesh_seed, r.source_seed, r.precision = 0.05, seed, seed + 1, fa.FP64_PC32
        return r
    mix = [req(2, 1, 10, 1), req(3, 1, 4, 2), req(4, 2, 2, 3)]
    runs = []
    for mode in (1, 3, 5):  # 1024, 4096 and 16384 wavefronts
        fa.tune_mode(mode)
        runs.append([(r.iters, r.residual, r.u_checksum) for r in fa.batch_run(mix)])
    assert runs[0] == runs[1] == runs[2]


def test_tuned_mode_fills_the_table():
    fa.tune_mode(0)
    _, _, gcx, geo, _, _ = both(3, 6, 0.05, seed=1)
    fa.WhitneyComplex(geo).mixed_kkt(1)
    lines = fa.tune_table()
    hits = [ln for ln in lines if "elmat_kkt d=3 k=1" in ln]
    assert hits, lines
    # the chosen candidate is the fastest timed one
    fields = hits[0].split()
    choice = int(fields[fields.index("choice") + 1])
    times = [float(t) for t in fields[fields.index("ms") + 1:]]
    assert len(times) == 2 and times[choice] == min(times)
    with pytest.raises(fa.FeecError):
        fa.tune_mode(65)
