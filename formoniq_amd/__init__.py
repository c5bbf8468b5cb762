"""formoniq_amd: formoniq's FEEC assembly-and-solve path on MI355X.

A host-side mirror of the reference's API for this path (luiswirth/formoniq,
paths relative to crates/), over the C ABI of libfeec (include/feec.h):

* ``CartesianGrid``      regge/src/mesher/cartesian.rs:109-233
* ``Complex``            simplicial/src/topology/complex.rs:278-336
* ``MeshLengthsSq``      regge/src/lengths/mesh.rs, regge/src/coord/mesh.rs:208-216
* ``WhitneyComplex``     formoniq/src/whitney_complex.rs:55-277, 321-465, 523-714
                         (``mass``, ``dif_trial``, ``dif_test``, ``dif_both``,
                         ``hdif_gram``, ``dif``, ``ndofs``, ``element``)
* ``solve_source``       formoniq/src/problems/elliptic.rs:132-182
* ``HodgeDirac``         formoniq/src/problems/dirac.rs:181-275
* ``solve_dirac_source`` formoniq/src/problems/dirac.rs:473-534

Everything computes on the GPU through libfeec; the module raises if the
library is not built.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import FeecError, Report, Request, Result, ServerOpts, SolveOpts, check, load

MASS, DIF_TRIAL, DIF_TEST, DIF_BOTH, HDIF_GRAM = 0, 1, 2, 3, 4
FULL, RELATIVE = 0, 1
EUCLIDEAN, MINKOWSKI = 0, 1
# feec_precision: FP64; FP32 Krylov inside FP64 refinement; FP64 Krylov with the
# polynomial in FP32 on the FP16 SELL copy
FP64, MIXED, FP64_PC32 = 0, 1, 2
# feec_pc_storage: how the preconditioner matrix the polynomial runs on is stored
PC_SELL, PC_CSR_HALF, PC_CSR = 0, 1, 2
# feec_preconditioner: the block preconditioner of the mixed KKT system
PC_AUTO, PC_CHEBYSHEV, PC_MULTIGRID = 0, 1, 2

__all__ = [
    "Context", "CartesianGrid", "Complex", "MeshLengthsSq", "WhitneyComplex", "Csr",
    "HodgeDirac", "Problem", "solve_source", "solve_dirac_source", "dirac_leapfrog", "dirac_irk",
    "solve_opts",
    "FeecError",
    "normal_vector", "MASS", "DIF_TRIAL", "DIF_TEST", "DIF_BOTH", "HDIF_GRAM", "FULL",
    "RELATIVE", "EUCLIDEAN", "MINKOWSKI", "FP64", "MIXED", "FP64_PC32", "Report", "Request", "Result",
    "ServerOpts", "SolveOpts", "Server", "batch_run", "load_vector", "PC_AUTO", "PC_CHEBYSHEV",
    "PC_MULTIGRID", "FRAME_REFERENCE", "FRAME_AMBIENT", "SimplexQuadRule",
    "cell_quadrature_points", "simplex_quadrature_points", "SourceForm", "source_load",
    "l2_projection", "whitney_at_nodes", "fe_l2_error", "derham_map", "DeviceArray",
    "gmsh2coord_cells", "load_cbor_cells", "load_cbor_lengths", "save_cbor_cells",
    "save_cbor_lengths", "tune_mode", "tune_table",
]


def _ptr(a, t):
    return a.ctypes.data_as(t)


def solve_opts(**kw) -> SolveOpts:
    o = SolveOpts()
    load().feec_solve_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


class Context:
    """One device and one HIP stream (feec_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(load().feec_ctx_create(device, C.byref(h)))
        self._h = h
        self.device = device

    def close(self, _load=load):
        if getattr(self, "_h", None):
            _load().feec_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def synchronize(self):
        check(load().feec_ctx_synchronize(self._h))

    def memory(self) -> int:
        b = C.c_int64()
        check(load().feec_ctx_memory(self._h, C.byref(b)))
        return b.value

    def allocations(self) -> int:
        """Device allocations made by this context so far (feec_ctx_allocations)."""
        c = C.c_int64()
        check(load().feec_ctx_allocations(self._h, C.byref(c)))
        return c.value

    def stream_rates(self, nbytes: int = 1 << 30):
        """Sustained (write, read, copy) GB/s of plain streaming kernels on this
        device (feec_ctx_stream_rates)."""
        r = np.zeros(3)
        check(load().feec_ctx_stream_rates(self._h, int(nbytes), _ptr(r, _lib.f64p)))
        return {"write_gbs": r[0], "read_gbs": r[1], "copy_gbs": r[2]}


_DEFAULT_CTX = None


def default_context() -> Context:
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = Context(0)
    return _DEFAULT_CTX


class CartesianGrid:
    """Unit box grid of n^dim cubes, Kuhn-triangulated (CartesianGrid::new_unit)."""

    def __init__(self, dim: int, n: int):
        self.dim, self.n = dim, n

    @staticmethod
    def new_unit(dim, n):
        return CartesianGrid(dim, n)

    def counts(self):
        nc, nv = C.c_int64(), C.c_int64()
        check(load().feec_kuhn_counts(self.dim, self.n, C.byref(nc), C.byref(nv)))
        return nc.value, nv.value

    def triangulate(self):
        """(cells [N_d, dim+1] int64 in generation order, coords [V, dim])."""
        nc, nv = self.counts()
        cells = np.zeros((nc, self.dim + 1), np.int64)
        coords = np.zeros((nv, self.dim), np.float64)
        check(load().feec_kuhn_grid(self.dim, self.n, _ptr(cells, _lib.i64p), _ptr(coords, _lib.f64p)))
        return cells, coords

    def jitter(self, coords, amplitude, seed, axes=None):
        """Interior-vertex jitter amplitude*h*(2U-1) per axis (SplitMix64(seed))."""
        coords = np.ascontiguousarray(coords, dtype=np.float64).copy()
        mask = 0
        for a in (range(self.dim) if axes is None else axes):
            mask |= 1 << a
        check(load().feec_jitter_interior(self.dim, self.n, _ptr(coords, _lib.f64p), amplitude,
                                          seed, mask))
        return coords


def load_vector(geometry: "MeshLengthsSq", grade, cochain):
    """The Galerkin load b = M_k c of a cochain (feec_load_vector), the
    type-correct source for solve_source (galerkin.rs:324-371)."""
    c = np.ascontiguousarray(cochain, dtype=np.float64)
    out = np.zeros_like(c)
    check(load().feec_load_vector(geometry._h, grade, _ptr(c, _lib.f64p), _ptr(out, _lib.f64p)))
    return out


# ---- analytic data: quadrature loads, L2 projection and error, de Rham map ------------
# feec_frame: components in each cell's reference frame, or in the mesh's coordinates
FRAME_REFERENCE, FRAME_AMBIENT = 0, 1


def _frame(frame):
    if frame in ("reference", FRAME_REFERENCE):
        return FRAME_REFERENCE
    if frame in ("ambient", FRAME_AMBIENT):
        return FRAME_AMBIENT
    raise ValueError(f"frame must be 'reference' or 'ambient', not {frame!r}")


def _array_in(a, shape):
    """(address, on_device, keepalive) of an input array of the given shape: a numpy
    array (host) or an object with data_ptr() and is_cuda (a torch tensor on the
    context's device). The library reads exactly that many doubles."""
    if isinstance(a, DeviceArray):
        if a.shape != tuple(shape):
            raise ValueError(f"expected shape {tuple(shape)}, got {a.shape}")
        return C.c_void_p(a.data_ptr()), 1, a
    if hasattr(a, "data_ptr") and getattr(a, "is_cuda", False):
        if not a.is_contiguous() or str(a.dtype) != "torch.float64":
            raise ValueError("device arrays must be contiguous float64")
        if tuple(a.shape) != tuple(shape):
            raise ValueError(f"expected shape {tuple(shape)}, got {tuple(a.shape)}")
        return C.c_void_p(a.data_ptr()), 1, a
    h = np.ascontiguousarray(a, dtype=np.float64)
    if h.shape != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {h.shape}")
    return C.c_void_p(h.ctypes.data), 0, h


def _cell_values_shape(geometry, grade, degree):
    d = geometry.complex.dim
    if not 0 <= grade <= d:
        raise FeecError(1, "grade out of range")
    return (geometry.complex.nsimplices(d), _npoints(d, degree), math.comb(d, grade))


_HIP = None


def _hip_runtime():
    """The HIP runtime libfeec links (libamdhip64), for DeviceArray."""
    global _HIP
    if _HIP is None:
        load()
        h = C.CDLL("libamdhip64.so", mode=C.RTLD_GLOBAL)
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipFree.argtypes = [C.c_void_p]
        _HIP = h
    return _HIP


class DeviceArray:
    """A float64 array in device memory (hipMalloc on the current device), for the
    on_device entry points; torch tensors on the device are accepted as well."""
    is_cuda = True

    def __init__(self, shape):
        self.shape = tuple(int(s) for s in shape)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self._p = C.c_void_p()
        hip = _hip_runtime()
        if hip.hipMalloc(C.byref(self._p), max(8 * self.size, 8)) != 0:
            raise FeecError(3, "hipMalloc failed")

    @staticmethod
    def from_numpy(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = DeviceArray(a.shape)
        if a.size and _hip_runtime().hipMemcpy(d._p, a.ctypes.data, a.nbytes, 1) != 0:
            raise FeecError(2, "hipMemcpy failed")
        return d

    def numpy(self):
        out = np.zeros(self.shape)
        if out.size and _hip_runtime().hipMemcpy(out.ctypes.data, self._p, out.nbytes, 2) != 0:
            raise FeecError(2, "hipMemcpy failed")
        return out

    def data_ptr(self):
        return self._p.value or 0

    def __del__(self):
        if getattr(self, "_p", None) and self._p.value and _HIP is not None:
            _HIP.hipFree(self._p)
            self._p = C.c_void_p()


def _array_out(like_device, shape):
    """An output array of the given shape on the same side as the input."""
    if like_device:
        d = DeviceArray(shape)
        return d, C.c_void_p(d.data_ptr())
    h = np.zeros(shape)
    return h, C.c_void_p(h.ctypes.data)


class SimplexQuadRule:
    """SimplexQuadRule::degree (simplicial/src/atlas/quadrature.rs:37-83): the
    Grundmann-Moller rule exact for polynomials of the given degree on the
    dim-simplex; `points` barycentric [np, dim+1], `weights` summing to one."""

    def __init__(self, dim, degree):
        n = C.c_int()
        check(load().feec_quadrature_rule(dim, degree, C.byref(n), None, None))
        self.dim, self.degree_ = dim, degree
        self.points = np.zeros((n.value, dim + 1))
        self.weights = np.zeros(n.value)
        check(load().feec_quadrature_rule(dim, degree, C.byref(n), _ptr(self.points, _lib.f64p),
                                          _ptr(self.weights, _lib.f64p)))

    @staticmethod
    def degree(dim, degree):
        return SimplexQuadRule(dim, degree)

    def npoints(self):
        return len(self.weights)


def _npoints(dim, degree):
    n = C.c_int()
    check(load().feec_quadrature_rule(dim, degree, C.byref(n), None, None))
    return n.value


def cell_quadrature_points(geometry: "MeshLengthsSq", degree, device=False):
    """Positions of every cell's quadrature nodes [N_d, np, dim] (the MeshPoints an
    integrand sees); needs a geometry made from coordinates."""
    d, nc = geometry.complex.dim, geometry.complex.nsimplices(geometry.complex.dim)
    out, addr = _array_out(device, (nc, _npoints(d, degree), d))
    check(load().feec_cell_quadrature_points(geometry._h, degree, addr, int(device)))
    return out


def simplex_quadrature_points(geometry: "MeshLengthsSq", grade, degree, device=False):
    """Positions of the quadrature nodes on every grade-k simplex [N_k, np_k, dim]."""
    d, ns = geometry.complex.dim, geometry.complex.nsimplices(grade)
    out, addr = _array_out(device, (ns, _npoints(grade, degree), d))
    check(load().feec_simplex_quadrature_points(geometry._h, grade, degree, addr, int(device)))
    return out


def source_load(geometry: "MeshLengthsSq", grade, values, degree=1, frame="reference"):
    """SourceForm(...).assemble (operators.rs:599-635, galerkin.rs:205-312): the load
    vector of a k-form given at every cell's degree-`degree` nodes, values
    [N_d, np, C(d, k)] (numpy, or a float64 torch tensor on the device)."""
    addr, dev, keep = _array_in(values, _cell_values_shape(geometry, grade, degree))
    out, oaddr = _array_out(dev, (geometry.complex.nsimplices(grade),))
    check(load().feec_source_load(geometry._h, grade, degree, _frame(frame), addr, oaddr, dev))
    return out


class SourceForm:
    """SourceForm::new(source, Some(qr)) (operators.rs:599-635) for a field given by
    its values at the cell nodes of the degree-`degree` rule."""

    def __init__(self, values, grade, degree=1, frame="reference"):
        self.values, self.grade, self.degree, self.frame = values, grade, degree, frame

    def assemble(self, geometry: "MeshLengthsSq"):
        return source_load(geometry, self.grade, self.values, self.degree, self.frame)


def l2_projection(geometry: "MeshLengthsSq", grade, values, degree=1, frame="reference",
                  mass: "Csr" = None):
    """l2_projection (fe.rs:87-123): (cochain, Report) of the mass solve M_k c = load.
    mass: the grade-k mass (WhitneyComplex(geometry).mass(grade)) to reuse, or None
    to assemble it inside the call."""
    addr, dev, keep = _array_in(values, _cell_values_shape(geometry, grade, degree))
    out, oaddr = _array_out(dev, (geometry.complex.nsimplices(grade),))
    rep = Report()
    check(load().feec_l2_projection(geometry._h, grade, degree, _frame(frame), addr,
                                    mass._h if mass is not None else None, oaddr, C.byref(rep),
                                    dev))
    return out, rep


def whitney_at_nodes(geometry: "MeshLengthsSq", grade, cochain, degree=3):
    """WhitneyInterpolant::eval (interpolant.rs:55-62) at every cell's nodes, in the
    cells' reference frames: [N_d, np, C(d, k)]."""
    shape = _cell_values_shape(geometry, grade, degree)
    addr, dev, keep = _array_in(cochain, (geometry.complex.nsimplices(grade),))
    out, oaddr = _array_out(dev, shape)
    check(load().feec_whitney_at_nodes(geometry._h, grade, degree, addr, oaddr, dev))
    return out


def fe_l2_error(cochain, exact, geometry: "MeshLengthsSq", grade, frame="reference"):
    """fe_l2_error (fe.rs:50-74): |exact - W c|_(L^2) with exact given at the
    degree-3 cell nodes."""
    shape = _cell_values_shape(geometry, grade, 3)
    caddr, cdev, ckeep = _array_in(cochain, (geometry.complex.nsimplices(grade),))
    eaddr, edev, ekeep = _array_in(exact, shape)
    if cdev != edev:
        raise ValueError("cochain and exact values must both be host or both device arrays")
    err = C.c_double()
    check(load().feec_l2_error(geometry._h, grade, caddr, _frame(frame), eaddr, C.byref(err), cdev))
    return err.value


def derham_map(geometry: "MeshLengthsSq", grade, values, degree=1):
    """derham_map (project.rs:55-96): the cochain of integrals of an ambient k-form
    given at the degree-`degree` nodes of every k-simplex [N_k, np_k, C(d, k)]."""
    d = geometry.complex.dim
    if not 0 <= grade <= d:
        raise FeecError(1, "grade out of range")
    ns = geometry.complex.nsimplices(grade)
    addr, dev, keep = _array_in(values, (ns, _npoints(grade, degree), math.comb(d, grade)))
    out, oaddr = _array_out(dev, (ns,))
    check(load().feec_derham_map(geometry._h, grade, degree, addr, oaddr, dev))
    return out


# ---- request ingestion: gmsh and CBOR ------------------------------------------------
def gmsh2coord_cells(data: bytes):
    """gmsh2coord_cells_ordered (regge/src/io/gmsh.rs:84-175) on MSH 4.1 ASCII bytes:
    (dim, cells [ncells, dim+1] as the file words them, renumbered; coords [nverts, 2 or 3])."""
    data = bytes(data)
    d, cd, nv, nc = C.c_int(), C.c_int(), C.c_int64(), C.c_int64()
    check(load().feec_gmsh_read(data, len(data), C.byref(d), C.byref(cd), C.byref(nv),
                                C.byref(nc), None, None))
    cells = np.zeros((nc.value, d.value + 1), np.int64)
    coords = np.zeros((nv.value, cd.value))
    check(load().feec_gmsh_read(data, len(data), C.byref(d), C.byref(cd), C.byref(nv),
                                C.byref(nc), _ptr(cells, _lib.i64p), _ptr(coords, _lib.f64p)))
    return d.value, cells, coords


def load_cbor_cells(data: bytes):
    """Complex::load (complex.rs:91-99) from CBOR bytes: (dim, cells [ncells, dim+1])."""
    data = bytes(data)
    d, nc = C.c_int(), C.c_int64()
    check(load().feec_cbor_read_cells(data, len(data), C.byref(d), C.byref(nc), None))
    cells = np.zeros((nc.value, d.value + 1), np.int64)
    check(load().feec_cbor_read_cells(data, len(data), C.byref(d), C.byref(nc),
                                      _ptr(cells, _lib.i64p)))
    return d.value, cells


def load_cbor_lengths(data: bytes):
    """MeshLengthsSq::load (lengths/mesh.rs:244-251) from CBOR bytes: the squared lengths."""
    data = bytes(data)
    n = C.c_int64()
    check(load().feec_cbor_read_lengths(data, len(data), C.byref(n), None))
    out = np.zeros(n.value)
    check(load().feec_cbor_read_lengths(data, len(data), C.byref(n), _ptr(out, _lib.f64p)))
    return out


def _cbor_out(fn, *args):
    n = C.c_int64(0)
    check(fn(*args, None, C.byref(n)))
    buf = C.create_string_buffer(n.value)
    check(fn(*args, buf, C.byref(n)))
    return buf.raw[: n.value]


def save_cbor_cells(cells):
    """Complex::save: the CBOR of the cell skeleton (cells [ncells, dim+1])."""
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    return _cbor_out(load().feec_cbor_write_cells, cells.shape[1] - 1, cells.shape[0],
                     _ptr(cells, _lib.i64p))


def save_cbor_lengths(lengths_sq):
    """MeshLengthsSq::save: the CBOR of a squared-length vector."""
    v = np.ascontiguousarray(lengths_sq, dtype=np.float64)
    return _cbor_out(load().feec_cbor_write_lengths, v.size, _ptr(v, _lib.f64p))


# ---- autotuning ----------------------------------------------------------------------
def tune_mode(mode: int):
    """0: tune bit-neutral launch choices by timing (default); c >= 1: always take
    candidate c-1 (the element kernels: 1 fused, 2 split by grade)."""
    check(load().feec_tune_mode(int(mode)))


def tune_table():
    """The per-device table of tuned launch choices: a list of text lines."""
    n = C.c_int64(0)
    check(load().feec_tune_table(None, C.byref(n)))
    buf = C.create_string_buffer(n.value)
    check(load().feec_tune_table(buf, C.byref(n)))
    return [line for line in buf.value.decode().splitlines() if line]


def normal_vector(n, seed):
    out = np.zeros(n)
    check(load().feec_normal_vector(n, seed, _ptr(out, _lib.f64p)))
    return out


class Csr:
    """A device CSR (feec_csr) and its host copy."""

    def __init__(self, handle, owner):
        self._h = handle
        self._owner = owner  # keeps the context that owns the buffers alive
        nr, nc, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        check(load().feec_csr_shape(handle, C.byref(nr), C.byref(nc), C.byref(nnz)))
        self.shape = (nr.value, nc.value)
        self.nnz = nnz.value
        self._host = None

    def __del__(self, _load=load):
        # _load is bound at definition: module globals are gone at interpreter exit
        if getattr(self, "_h", None):
            _load().feec_csr_destroy(self._h)
            self._h = None

    def host(self):
        if self._host is None:
            rp = np.zeros(self.shape[0] + 1, np.int64)
            col = np.zeros(max(self.nnz, 1), np.int32)
            val = np.zeros(max(self.nnz, 1), np.float64)
            check(load().feec_csr_copy(self._h, _ptr(rp, _lib.i64p), _ptr(col, _lib.i32p),
                                       _ptr(val, _lib.f64p)))
            self._host = (rp, col[: self.nnz], val[: self.nnz])
        return self._host

    @property
    def row_ptr(self):
        return self.host()[0]

    @property
    def col(self):
        return self.host()[1]

    @property
    def val(self):
        return self.host()[2]

    def to_scipy(self):
        import scipy.sparse as sp
        rp, col, val = self.host()
        return sp.csr_matrix((val, col, rp), shape=self.shape)

    def __matmul__(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.shape[0])
        check(load().feec_csr_spmv(self._h, _ptr(x, _lib.f64p), _ptr(y, _lib.f64p)))
        return y

    def cg(self, b, rtol=1e-10, max_iters=10_000, raise_on_failure=True):
        """krylov::cg with Jacobi (krylov.rs:48-95) on the device: (x, Report)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.shape[0])
        rep = Report()
        st = load().feec_csr_cg(self._h, _ptr(b, _lib.f64p), _ptr(x, _lib.f64p), float(rtol),
                                int(max_iters), C.byref(rep))
        if raise_on_failure:
            check(st)
        return x, rep

    def view(self, on_device):
        """feec_csr_view: (nrows, ncols, nnz, row_ptr, col, val) with the three
        arrays as raw addresses (device memory when on_device)."""
        nr, nc, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        rp, col, val = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(load().feec_csr_view(self._h, C.byref(nr), C.byref(nc), C.byref(nnz), C.byref(rp),
                                   C.byref(col), C.byref(val), int(bool(on_device))))
        return nr.value, nc.value, nnz.value, rp.value, col.value, val.value

    def spmv_device(self, x_addr, y_addr):
        """y = A x on device arrays (feec_csr_spmv_device)."""
        check(load().feec_csr_spmv_device(self._h, C.c_void_p(x_addr), C.c_void_p(y_addr)))

    def cg_noalloc(self, b, rtol=1e-10, max_iters=10_000, raise_on_failure=True):
        """cg() on the library's fixed device scratch (feec_csr_cg_noalloc):
        no device allocation, at most 262144 rows."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.shape[0])
        rep = Report()
        st = load().feec_csr_cg_noalloc(self._h, _ptr(b, _lib.f64p), _ptr(x, _lib.f64p), float(rtol),
                                        int(max_iters), C.byref(rep))
        if raise_on_failure:
            check(st)
        return x, rep


class Complex:
    """A simplicial complex numbered on the device (Complex::from_cells)."""

    def __init__(self, dim, cells, nverts=None, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        cells = np.ascontiguousarray(cells, dtype=np.int64)
        if nverts is None:
            nverts = int(cells.max()) + 1
        h = C.c_void_p()
        check(load().feec_mesh_create(self.ctx._h, dim, nverts, cells.shape[0], _ptr(cells, _lib.i64p),
                                      C.byref(h)))
        self._h = h
        self.dim = dim
        cnt = np.zeros(dim + 1, np.int64)
        check(load().feec_mesh_counts(h, _ptr(cnt, _lib.i64p)))
        self.counts = cnt

    def __del__(self, _load=load):
        # _load is bound at definition: module globals are gone at interpreter exit
        if getattr(self, "_h", None):
            _load().feec_mesh_destroy(self._h)
            self._h = None

    def nsimplices(self, k):
        return int(self.counts[k]) if 0 <= k <= self.dim else 0

    def skeleton(self, k):
        out = np.zeros((self.nsimplices(k), k + 1), np.int64)
        check(load().feec_mesh_skeleton(self._h, k, _ptr(out, _lib.i64p)))
        return out

    def cells(self):
        return self.skeleton(self.dim)

    def cell_faces(self, k):
        from math import comb
        out = np.zeros((self.nsimplices(self.dim), comb(self.dim + 1, k + 1)), np.int64)
        check(load().feec_mesh_cell_faces(self._h, k, _ptr(out, _lib.i64p)))
        return out

    def boundary_flags(self, k):
        out = np.zeros(self.nsimplices(k), np.uint8)
        check(load().feec_mesh_boundary(self._h, k, _ptr(out, _lib.u8p)))
        return out

    def orientation(self):
        out = np.zeros(self.nsimplices(self.dim), np.int8)
        check(load().feec_mesh_orientation(self._h, _ptr(out, _lib.i8p)))
        return out

    def coboundary(self, k, bc=FULL):
        h = C.c_void_p()
        check(load().feec_mesh_coboundary(self._h, bc, k, C.byref(h)))
        return Csr(h, self)

    def _ndofs(self, k, bc):
        if bc == RELATIVE:
            return int((self.boundary_flags(k) == 0).sum())
        return self.nsimplices(k)

    def betti_number(self, k, bc=FULL):
        """b_k (Complex::betti_number), or relative to the boundary with
        bc=RELATIVE (relative_betti_number), from exact ranks."""
        b = C.c_int64()
        check(load().feec_betti_number(self._h, bc, k, C.byref(b)))
        return int(b.value)

    def _generators(self, fn, k, bc):
        cnt = C.c_int64()
        check(fn(self._h, bc, k, C.byref(cnt), None))
        out = np.zeros((cnt.value, self._ndofs(k, bc)), np.int64)
        if cnt.value:
            check(fn(self._h, bc, k, C.byref(cnt), _ptr(out, _lib.i64p)))
        return out

    def cohomology_generators(self, k, bc=FULL):
        """Integral cocycles of a basis of the free part of H^k, one per row, on
        the complex's DOFs (relative: interior numbering)."""
        return self._generators(load().feec_cohomology_generators, k, bc)

    def homology_generators(self, k, bc=FULL):
        """Integral cycles of a basis of the free part of H_k, one per row."""
        return self._generators(load().feec_homology_generators, k, bc)


class MeshLengthsSq:
    """Signed squared edge lengths on the device (a Regge metric)."""

    def __init__(self, complex_: Complex, handle):
        self.complex = complex_
        self._h = handle

    @staticmethod
    def from_lengths(complex_: Complex, lengths):
        lengths = np.ascontiguousarray(lengths, dtype=np.float64)
        h = C.c_void_p()
        check(load().feec_geom_from_lengths(complex_._h, _ptr(lengths, _lib.f64p), C.byref(h)))
        return MeshLengthsSq(complex_, h)

    @staticmethod
    def from_coords(complex_: Complex, coords, minkowski=False):
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        h = C.c_void_p()
        check(load().feec_geom_from_coords(complex_._h, _ptr(coords, _lib.f64p),
                                           MINKOWSKI if minkowski else EUCLIDEAN, C.byref(h)))
        return MeshLengthsSq(complex_, h)

    def __del__(self, _load=load):
        # _load is bound at definition: module globals are gone at interpreter exit
        if getattr(self, "_h", None):
            _load().feec_geom_destroy(self._h)
            self._h = None

    def lengths(self):
        out = np.zeros(self.complex.nsimplices(1))
        check(load().feec_geom_lengths(self._h, _ptr(out, _lib.f64p)))
        return out

    def cell_metrics(self):
        d, nc = self.complex.dim, self.complex.nsimplices(self.complex.dim)
        g, gi, det = np.zeros((nc, d, d)), np.zeros((nc, d, d)), np.zeros(nc)
        check(load().feec_cell_metrics(self._h, _ptr(g, _lib.f64p), _ptr(gi, _lib.f64p),
                                       _ptr(det, _lib.f64p)))
        return g, gi, det

    def dirac_element_time(self, reps=10):
        """(seconds per launch, values per cell) of the all-grade Hodge-Dirac element kernel."""
        sec, per = C.c_double(), C.c_int64()
        check(load().feec_dirac_element_time(self._h, reps, C.byref(sec), C.byref(per)))
        return sec.value, per.value

    def causal_census(self):
        counts = np.zeros(3, np.int64)
        riem = np.zeros(self.complex.nsimplices(self.complex.dim), np.uint8)
        check(load().feec_causal_census(self._h, _ptr(counts, _lib.i64p), _ptr(riem, _lib.u8p)))
        return counts, riem


def element_shape(dim, kind, grade):
    r, c = C.c_int(), C.c_int()
    check(load().feec_element_shape(dim, kind, grade, C.byref(r), C.byref(c)))
    return r.value, c.value


class WhitneyComplex:
    """HilbertComplex over the Whitney forms of a mesh, full or relative."""

    def __init__(self, geometry: MeshLengthsSq, relative: bool = False):
        self.geometry = geometry
        self.topology = geometry.complex
        self.relative = relative
        self.bc = RELATIVE if relative else FULL
        self.dim = self.topology.dim

    def ndofs(self, k):
        if not 0 <= k <= self.dim:
            return 0
        if self.relative:
            return int((self.topology.boundary_flags(k) == 0).sum())
        return self.topology.nsimplices(k)

    def element(self, kind, grade):
        r, c = element_shape(self.dim, kind, grade)
        out = np.zeros((self.topology.nsimplices(self.dim), r, c))
        if r * c:
            check(load().feec_element_matrices(self.geometry._h, kind, grade, _ptr(out, _lib.f64p)))
        return out

    def assemble(self, kind, grade) -> Csr:
        h = C.c_void_p()
        check(load().feec_assemble(self.geometry._h, self.bc, kind, grade, C.byref(h)))
        return Csr(h, self)

    def mass(self, k):
        return self.assemble(MASS, k)

    def dif_trial(self, k):
        return self.assemble(DIF_TRIAL, k)

    def dif_test(self, k):
        return self.assemble(DIF_TEST, k)

    def dif_both(self, k):
        return self.assemble(DIF_BOTH, k)

    def hdif_gram(self, k):
        return self.assemble(HDIF_GRAM, k)

    def dif(self, k):
        return self.topology.coboundary(k, self.bc)

    def _harmonic_reading(self, fn, k):
        cnt = C.c_int64()
        check(fn(self.geometry._h, self.bc, k, C.byref(cnt), None))
        n = self.ndofs(k)
        out = np.zeros((max(cnt.value, 1), n))
        if cnt.value:
            check(fn(self.geometry._h, self.bc, k, C.byref(cnt), _ptr(out, _lib.f64p)))
        return out[: cnt.value].T.copy()

    def harmonics(self, k):
        """An M_k-orthonormal harmonic basis, ndofs x b_k (solve_harmonics)."""
        return self._harmonic_reading(load().feec_harmonics, k)

    def harmonics_integral(self, k):
        """The period-normalized harmonic basis, ndofs x b_k (Harmonics::integral)."""
        return self._harmonic_reading(load().feec_harmonics_integral, k)

    def mixed_kkt(self, k):
        h = C.c_void_p()
        sizes = np.zeros(3, np.int64)
        check(load().feec_mixed_kkt(self.geometry._h, self.bc, k, C.byref(h), _ptr(sizes, _lib.i64p)))
        return Csr(h, self), tuple(int(s) for s in sizes)


def solve_source(complex_: WhitneyComplex, source_galvec, grade, opts: SolveOpts | None = None,
                 raise_on_failure=True):
    """elliptic::solve_source: (sigma, u, p, report) with ambient cochains."""
    topo = complex_.topology
    src = np.ascontiguousarray(source_galvec, dtype=np.float64)
    assert src.shape == (topo.nsimplices(grade),)
    sigma = np.zeros(max(topo.nsimplices(grade - 1), 1))
    u = np.zeros(topo.nsimplices(grade))
    p = np.zeros(8)
    rep = Report()
    st = load().feec_solve_source(complex_.geometry._h, complex_.bc, grade, _ptr(src, _lib.f64p),
                                  C.byref(opts) if opts is not None else None, _ptr(sigma, _lib.f64p),
                                  _ptr(u, _lib.f64p), _ptr(p, _lib.f64p), C.byref(rep))
    if raise_on_failure:
        check(st)
    nh = topo.betti_number(grade, complex_.bc)
    return sigma[: topo.nsimplices(grade - 1)], u, p[:nh], rep


class Problem:
    """A device-resident mixed Hodge-Laplace problem (pattern and Krylov state kept)."""

    def __init__(self, complex_: WhitneyComplex, grade, opts: SolveOpts | None = None):
        self.complex = complex_
        self.grade = grade
        self.opts = opts if opts is not None else solve_opts()
        h = C.c_void_p()
        check(load().feec_problem_create(complex_.geometry._h, complex_.bc, grade,
                                         C.byref(opts) if opts is not None else None, C.byref(h)))
        self._h = h
        sizes = np.zeros(3, np.int64)
        nnz = C.c_int64()
        check(load().feec_problem_sizes(h, _ptr(sizes, _lib.i64p), C.byref(nnz)))
        self.sizes = tuple(int(s) for s in sizes)
        self.nnz = nnz.value

    def __del__(self, _load=load):
        # _load is bound at definition: module globals are gone at interpreter exit
        if getattr(self, "_h", None):
            _load().feec_problem_destroy(self._h)
            self._h = None

    def assemble(self):
        check(load().feec_problem_assemble(self._h))

    def kernel_times(self, reps=20):
        """Seconds per launch: element kernel, KKT scatter, FP32 and FP64 MINRES
        SpMV, FP32 and FP64 Chebyshev step."""
        out = np.zeros(6)
        check(load().feec_problem_kernel_times(self._h, reps, _ptr(out, _lib.f64p)))
        return dict(elmat=float(out[0]), scatter=float(out[1]), spmv32=float(out[2]),
                    spmv64=float(out[3]), cheb32=float(out[4]), cheb64=float(out[5]))

    def mg_kernel_times(self, reps=20, use_graph=True):
        """Seconds per launch with the multigrid preconditioner: finest-level
        sweep, block preconditioner application (FP64_PC32), product by the
        preconditioner matrix, one HX application B (graph replay, or launched
        one by one)."""
        out = np.zeros(4)
        check(load().feec_problem_mg_kernel_times(self._h, reps, 1 if use_graph else 0,
                                                  _ptr(out, _lib.f64p)))
        return dict(sweep_fine=float(out[0]), precond_apply=float(out[1]), pc_spmv32=float(out[2]),
                    hx_apply=float(out[3]))

    PC_FORMATS = {0: "csr_f32", 1: "csr_f16", 2: "sell64_f16", 3: "csr_f64", 4: "sell64_f64",
                  5: "sell64_f16_c32", 6: "sell64_f64_c32", 7: "multigrid_f32", 8: "multigrid_f64"}

    def apply_preconditioner(self, r, precision=None, storage=None):
        """z = B r (KKT order: sigma, u, harmonic rows) with the block preconditioner
        as a solve of `precision` with `storage` (PC_SELL, PC_CSR_HALF, PC_CSR)
        applies it (defaults: this problem's options)."""
        prec = self.opts.precision if precision is None else precision
        store = self.opts.precond_storage if storage is None else storage
        r = np.ascontiguousarray(r, dtype=np.float64)
        assert r.shape == (sum(self.sizes),)
        z = np.zeros_like(r)
        check(load().feec_problem_apply_preconditioner(self._h, prec, store, _ptr(r, _lib.f64p),
                                                       _ptr(z, _lib.f64p)))
        return z

    def preconditioner(self):
        """The block preconditioner as built: nonzeros, Chebyshev degree, lambda_max
        per block, bytes per stored CSR column index, the format a solve with this
        problem's options applies, the entries that format stores (SELL padding
        included) and the off-diagonal nonzeros."""
        nnz, deg, lam, cb = C.c_int64(), C.c_int32(), np.zeros(2), C.c_int32()
        fmt, stored, offd = C.c_int32(), C.c_int64(), C.c_int64()
        check(load().feec_problem_preconditioner(self._h, C.byref(nnz), C.byref(deg),
                                                 _ptr(lam, _lib.f64p), C.byref(cb), C.byref(fmt),
                                                 C.byref(stored), C.byref(offd)))
        return dict(nnz=nnz.value, degree=deg.value, lambda_max=lam, column_bytes=cb.value,
                    format=self.PC_FORMATS[fmt.value], stored=stored.value,
                    offdiag_nnz=offd.value)

    def upload_source(self, source):
        """Make `source` (N_k) resident on the device for solve()."""
        src = np.ascontiguousarray(source, dtype=np.float64)
        check(load().feec_problem_upload_source(self._h, _ptr(src, _lib.f64p)))

    def source_buffer(self):
        p = C.c_void_p()
        check(load().feec_problem_source_buffer(self._h, C.byref(p)))
        return p

    def solve(self, source=None) -> Report:
        rep = Report()
        if source is None:
            check(load().feec_problem_solve_device(self._h, self.source_buffer(), C.byref(rep)))
        else:
            src = np.ascontiguousarray(source, dtype=np.float64)
            check(load().feec_problem_solve(self._h, _ptr(src, _lib.f64p), C.byref(rep)))
        return rep

    def solve_cochain(self, cochain=None) -> Report:
        """Solve with the Galerkin load b = M_k c of a cochain c, formed on the
        device (feec_problem_solve_cochain*); None: c is in source_buffer()."""
        rep = Report()
        if cochain is None:
            check(load().feec_problem_solve_cochain_device(self._h, self.source_buffer(), C.byref(rep)))
        else:
            c = np.ascontiguousarray(cochain, dtype=np.float64)
            check(load().feec_problem_solve_cochain(self._h, _ptr(c, _lib.f64p), C.byref(rep)))
        return rep

    def solution(self):
        topo = self.complex.topology
        k = self.grade
        sigma = np.zeros(max(topo.nsimplices(k - 1), 1))
        u = np.zeros(topo.nsimplices(k))
        p = np.zeros(max(self.sizes[2], 1))
        check(load().feec_problem_solution(self._h, _ptr(sigma, _lib.f64p), _ptr(u, _lib.f64p),
                                           _ptr(p, _lib.f64p)))
        return sigma[: topo.nsimplices(k - 1)], u, p[: self.sizes[2]]


@dataclass
class HodgeDirac:
    offsets: np.ndarray
    mass_block: Csr
    op: Csr

    @staticmethod
    def assemble_signed(complex_: WhitneyComplex, sign):
        mb, op = C.c_void_p(), C.c_void_p()
        offs = np.zeros(complex_.dim + 2, np.int64)
        check(load().feec_hodge_dirac_assemble(complex_.geometry._h, complex_.bc, sign, C.byref(mb),
                                               C.byref(op), _ptr(offs, _lib.i64p)))
        return HodgeDirac(offs, Csr(mb, complex_), Csr(op, complex_))

    @staticmethod
    def assemble(complex_):
        return HodgeDirac.assemble_signed(complex_, -1.0)

    @staticmethod
    def assemble_selfadjoint(complex_):
        return HodgeDirac.assemble_signed(complex_, 1.0)

    def ndofs_total(self):
        return int(self.offsets[-1])


def dirac_leapfrog(complex_: WhitneyComplex, dt, nsteps, y0, keep_states=False):
    """solve_dirac_leapfrog's time loop (Leapfrog::step, time.rs:386-400) on the
    device: the flat field y0 of the complex (HodgeDirac offsets) after nsteps
    steps of dt, the staggered invariant of every state (nsteps + 1 values,
    time.rs:409-420) and the report (CG iterations of the mass solves); with
    keep_states, also every state as an (nsteps + 1, ndofs) array."""
    y = np.array(y0, dtype=np.float64, copy=True)
    energy = np.zeros(int(nsteps) + 1)
    states = np.zeros((int(nsteps) + 1, y.size)) if keep_states else None
    rep = Report()
    check(load().feec_dirac_leapfrog(complex_.geometry._h, complex_.bc, float(dt), int(nsteps),
                                     _ptr(y, _lib.f64p),
                                     _ptr(states, _lib.f64p) if keep_states else None,
                                     _ptr(energy, _lib.f64p), C.byref(rep)))
    return (y, energy, rep, states) if keep_states else (y, energy, rep)


def dirac_irk(complex_: WhitneyComplex, dt, nsteps, y0, stages=2, keep_states=False):
    """solve_dirac's time loop (LinearIrk::step with Tableau::gauss_legendre,
    time.rs:191-274) on the device: the flat field after nsteps steps, the
    energy 1/2 y^T M y of every state and the report (GMRES iterations)."""
    y = np.array(y0, dtype=np.float64, copy=True)
    energy = np.zeros(int(nsteps) + 1)
    states = np.zeros((int(nsteps) + 1, y.size)) if keep_states else None
    rep = Report()
    check(load().feec_dirac_irk(complex_.geometry._h, complex_.bc, int(stages), float(dt),
                                int(nsteps), _ptr(y, _lib.f64p),
                                _ptr(states, _lib.f64p) if keep_states else None,
                                _ptr(energy, _lib.f64p), C.byref(rep)))
    return (y, energy, rep, states) if keep_states else (y, energy, rep)


def solve_dirac_source(geometry: MeshLengthsSq, mass_term, load_field, lift, opts=None):
    """solve_dirac_source on the relative complex; flat full-layout fields
    (lift None: the zero lift)."""
    f = np.ascontiguousarray(load_field, dtype=np.float64)
    l = None if lift is None else np.ascontiguousarray(lift, dtype=np.float64)
    out = np.zeros_like(f)
    rep = Report()
    check(load().feec_solve_dirac_source(geometry._h, mass_term, _ptr(f, _lib.f64p),
                                         None if l is None else _ptr(l, _lib.f64p),
                                         C.byref(opts) if opts is not None else None,
                                         _ptr(out, _lib.f64p), C.byref(rep)))
    return out, rep


def batch_run(requests, opts: SolveOpts | None = None, ctx: Context | None = None):
    ctx = ctx or default_context()
    n = len(requests)
    reqs = (Request * n)(*requests)
    res = (Result * n)()
    check(load().feec_batch_run(ctx._h, n, reqs, C.byref(opts) if opts is not None else None, res))
    return list(res)


class Server:
    """Continuous-batching request loop (feec_server_*)."""

    def __init__(self, opts: ServerOpts, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        check(load().feec_server_create(self.ctx._h, C.byref(opts), C.byref(h)))
        self._h = h

    def submit(self, req: Request) -> int:
        t = C.c_int64()
        check(load().feec_server_submit(self._h, C.byref(req), C.byref(t)))
        return t.value

    def poll(self, ticket):
        r = Result()
        st = load().feec_server_poll(self._h, ticket, C.byref(r))
        if st == 10:
            return None
        check(st)
        return r

    def solution(self, ticket):
        n = C.c_int64()
        check(load().feec_server_solution(self._h, ticket, None, C.byref(n)))
        u = np.zeros(n.value)
        check(load().feec_server_solution(self._h, ticket, _ptr(u, _lib.f64p), C.byref(n)))
        return u

    def drain(self):
        check(load().feec_server_drain(self._h))

    def release(self, ticket):
        """Forget a finished request (feec_server_release)."""
        check(load().feec_server_release(self._h, ticket))

    def held(self):
        """(requests held, solution doubles held on the host)."""
        e, d = C.c_int64(), C.c_int64()
        check(load().feec_server_held(self._h, C.byref(e), C.byref(d)))
        return e.value, d.value

    def close(self, _load=load):
        if getattr(self, "_h", None):
            _load().feec_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
