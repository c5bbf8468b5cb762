"""The boundary pieces around the Galerkin matrices, on the GPU against the
oracle: the load vector b = M_k c (galerkin.rs:324-371), problem solves from a
cochain, the CSR view with device pointers and the device-vector operations
(the seam of iterative/src/lib.rs:58-186)."""
import ctypes as C

import numpy as np
import pytest

import formoniq_amd as fa
from oracle import oracle as o

from _common import both

pytestmark = pytest.mark.gpu


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = C.CDLL(name)
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipFree.argtypes = [C.c_void_p]
            return h
        except OSError:
            pass
    raise RuntimeError("HIP runtime not found")


class DeviceArray:
    """A raw device allocation for the test, filled and read through hipMemcpy."""

    def __init__(self, hip, host):
        self.hip, self.n = hip, host.size
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(8 * self.n, 8)) == 0
        assert hip.hipMemcpy(self.ptr, host.ctypes.data, 8 * self.n, 1) == 0

    def get(self, dtype=np.float64, count=None):
        out = np.zeros(self.n if count is None else count, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _read(hip, addr, count, dtype):
    out = np.zeros(count, dtype)
    if count:
        assert hip.hipMemcpy(out.ctypes.data, C.c_void_p(addr), out.nbytes, 2) == 0
    return out


@pytest.mark.parametrize("d,n,amp", [(2, 6, 0.1), (2, 4, 0.0), (3, 3, 0.05), (4, 2, 0.02)])
def test_load_vector_matches_oracle(d, n, amp):
    """b = M_k c bit for bit: the mass assembled as the oracle's, each row
    summed in column order (the oracle's sequential SpMV)."""
    ocx, oL, gcx, geo, _, _ = both(d, n, amp, seed=d + 20)
    oc = o.WhitneyComplex(ocx, oL)
    for k in range(d + 1):
        c = fa.normal_vector(ocx.nsimplices(k), 30 + k)
        b = fa.load_vector(geo, k, c)
        want = oc.mass(k) @ c
        assert np.array_equal(b, want), (k, int((b != want).sum()))


@pytest.mark.parametrize("precision", [fa.FP64, fa.FP64_PC32])
@pytest.mark.parametrize("rel", [False, True])
def test_cochain_solve_is_load_then_solve(precision, rel):
    """A problem solved from a cochain (b = M_k c formed on the device from
    its own element blocks) equals solve_source on feec_load_vector's b, bit
    for bit, and the KKT residual of the oracle's system with that b is
    certified."""
    ocx, oL, gcx, geo, _, _ = both(3, 4, 0.05, seed=3)
    wc = fa.WhitneyComplex(geo, relative=rel)
    opts = fa.solve_opts(precision=precision)
    c = fa.normal_vector(ocx.nsimplices(1), 8)
    b = fa.load_vector(geo, 1, c)
    prob = fa.Problem(wc, 1, opts)
    prob.assemble()
    r1 = prob.solve_cochain(c)
    s1, u1, _ = prob.solution()
    sig, u, p, r0 = fa.solve_source(wc, b, 1, opts)
    assert r1.converged and r1.iters == r0.iters
    np.testing.assert_array_equal(u1, u)
    np.testing.assert_array_equal(s1, sig)
    oc = o.WhitneyComplex(ocx, oL, relative=rel)
    kkt, rhs, ns, nu = o.assemble_mixed_kkt(oc, b, 1, oc.harmonics(1))
    x = np.concatenate([sig[oc.inclusion_indices(0)], u[oc.inclusion_indices(1)], p])
    assert np.linalg.norm(rhs - kkt.to_scipy() @ x) / np.linalg.norm(rhs) <= 1e-9


def test_csr_view_host_and_device():
    _, _, gcx, geo, _, _ = both(3, 3, 0.05, seed=4)
    A = fa.WhitneyComplex(geo).hdif_gram(1)
    rp, col, val = A.host()
    hip = _hip()
    nr, nc, nnz, drp, dcol, dval = A.view(on_device=True)
    assert (nr, nc, nnz) == (A.shape[0], A.shape[1], A.nnz)
    assert np.array_equal(_read(hip, drp, nr + 1, np.int64), rp)
    assert np.array_equal(_read(hip, dcol, nnz, np.int32), col)
    assert np.array_equal(_read(hip, dval, nnz, np.float64), val)
    _, _, _, hrp, hcol, hval = A.view(on_device=False)
    assert np.array_equal(np.ctypeslib.as_array((C.c_int64 * (nr + 1)).from_address(hrp)), rp)
    assert np.array_equal(np.ctypeslib.as_array((C.c_int32 * nnz).from_address(hcol)), col)
    assert np.array_equal(np.ctypeslib.as_array((C.c_double * nnz).from_address(hval)), val)


def test_device_vector_operations():
    """LinearOperator::apply and InnerProductSpace::{dot, add_scaled} on
    device arrays: the SpMV is the library's own kernel (equal to the host
    entry bit for bit); dot and axpy against numpy."""
    ctx = fa.default_context()
    _, _, gcx, geo, _, _ = both(3, 4, 0.05, seed=5)
    A = fa.WhitneyComplex(geo).hdif_gram(1)
    n = A.shape[0]
    hip = _hip()
    x = fa.normal_vector(n, 1)
    y = fa.normal_vector(n, 2)
    dx, dy = DeviceArray(hip, x), DeviceArray(hip, y)
    try:
        A.spmv_device(dx.ptr.value, dy.ptr.value)
        assert np.array_equal(dy.get(), (A @ x))
        out = C.c_double()
        fa.check(fa.load().feec_vec_dot_device(ctx._h, n, dx.ptr, dx.ptr, C.byref(out)))
        assert abs(out.value - x @ x) <= 1e-12 * (x @ x)
        fa.check(fa.load().feec_vec_axpy_device(ctx._h, n, 0.5, dx.ptr, dy.ptr))
        np.testing.assert_allclose(dy.get(), (A @ x) + 0.5 * x, rtol=1e-15, atol=1e-15)
    finally:
        dx.free()
        dy.free()
