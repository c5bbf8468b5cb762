"""ctypes face of the CPU FP64 oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: this module is the parity checker and the CPU
baseline. The tests, ``__graft_entry__.smoke()`` and the benchmark's
``cpu_baseline`` leg import it; the product package ``formoniq_amd`` never
does.

The classes mirror the reference's host API for the hot path so that tests read
like formoniq's own:

* ``Complex``            simplicial/src/topology/complex.rs:298-336
* ``WhitneyComplex``     formoniq/src/whitney_complex.rs:321-465 (full)
                         formoniq/src/whitney_complex.rs:523-714 (relative)
* ``HodgeBlocks``        formoniq/src/hodge.rs:62-99
* ``assemble_mixed_kkt`` formoniq/src/problems/elliptic.rs:54-114
* ``solve_source``       formoniq/src/problems/elliptic.rs:132-182
* ``hodge_dirac``        formoniq/src/problems/dirac.rs:213-258
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MASS, DIF_TRIAL, DIF_TEST, DIF_BOTH = 0, 1, 2, 3

# Deliberate mistakes in the KKT and Hodge-Dirac assembly below, switched on
# only by the tests that check the reference's laws catch them (0 = the
# restatement as written):
#   11 U_k = D_{k-1} M_{k-1} (the mass of the wrong grade, strong side)
#   12 U_k scaled by 2
#   13 the sub-diagonal block negated (the super-diagonal kept)
#   21 the sigma-row sign flip left off the coupling block
#   22 the sigma mass scaled by 2
#   23 the u-u block (dif_both(k+1)) negated
# and in the multigrid / auxiliary-space preconditioner:
#   31 the midpoint prolongation weights (1/2, 1/2) replaced by (3/4, 1/4)
#   32 the post-smoothing sweeps of the V-cycle left out
#   33 the vector-nodal factor 1/(k+1)! replaced by 1/k!
# and in the analytic-data functions:
#   41 the Grundmann-Moeller weight sign (-1)^i left off
#   42 the k! of the Whitney values left off
#   43 the load paired by the Euclidean product (Lambda^k g^-1 replaced by I)
#   44 the ambient pullback by Lambda^k(J^T) (J transposed)
#   45 the de Rham map's 1/k! (the unit simplex volume) left off
#   46 the L2 error without the cell volume
# and in the exact (co)homology and harmonic bases:
#   51 quotient_generators keeps the kernel vectors that do NOT enlarge the span
#   52 primitive() leaves the first nonzero entry negative
#   53 the period normalization left out (integral = H)
#   54 the harmonic projection h = z + D p (the sign of the exact part flipped)
MUTATION = 0

i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)


class _Csr(C.Structure):
    _fields_ = [("nrows", C.c_int64), ("ncols", C.c_int64), ("nnz", C.c_int64),
                ("row_ptr", i64p), ("col", i64p), ("val", f64p)]


class _Report(C.Structure):
    _fields_ = [("iters", C.c_int64), ("residual", C.c_double), ("converged", C.c_int)]


def lib(path=None):
    """The oracle library; `path` loads another build of the same source (the
    mutation test compiles deliberately broken variants)."""
    global _LIB
    if path is not None or _LIB is None:
        path = path or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        P = C.c_void_p
        sig = {
            "orc_binom": (C.c_int64, [C.c_int64, C.c_int64]),
            "orc_local_subsets": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int)]),
            "orc_form_shape": (None, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "orc_kuhn_ncells": (C.c_int64, [C.c_int, C.c_int]),
            "orc_kuhn_nverts": (C.c_int64, [C.c_int, C.c_int]),
            "orc_kuhn": (None, [C.c_int, C.c_int, i64p, f64p]),
            "orc_complex_new": (P, [C.c_int, C.c_int64, i64p]),
            "orc_complex_free": (None, [P]),
            "orc_nsimplices": (C.c_int64, [P, C.c_int]),
            "orc_skeleton": (None, [P, C.c_int, i64p]),
            "orc_cell_faces": (None, [P, C.c_int, i64p]),
            "orc_boundary_flags": (None, [P, C.c_int, u8p]),
            "orc_coboundary": (C.POINTER(_Csr), [P, C.c_int]),
            "orc_orientation": (C.c_int, [P, C.POINTER(C.c_int8)]),
            "orc_lengths_from_coords": (None, [P, f64p, C.c_int, f64p]),
            "orc_cell_metric": (None, [C.c_int, f64p, f64p, f64p, f64p]),
            "orc_element_matrix": (None, [C.c_int, C.c_int, C.c_int, f64p, f64p]),
            "orc_element_matrices": (None, [P, f64p, C.c_int, C.c_int, f64p, C.c_int]),
            "orc_assemble": (C.POINTER(_Csr), [P, f64p, C.c_int, C.c_int, C.c_int]),
            "orc_csr_free": (None, [C.POINTER(_Csr)]),
            "orc_csr_add": (C.POINTER(_Csr), [C.POINTER(_Csr), C.POINTER(_Csr)]),
            "orc_csr_mul": (C.POINTER(_Csr), [C.POINTER(_Csr), C.POINTER(_Csr)]),
            "orc_csr_transpose": (C.POINTER(_Csr), [C.POINTER(_Csr)]),
            "orc_csr_select": (C.POINTER(_Csr), [C.POINTER(_Csr), u8p, u8p]),
            "orc_csr_block": (C.POINTER(_Csr), [C.c_int, C.POINTER(C.POINTER(_Csr)), f64p, i64p, i64p]),
            "orc_spmv": (None, [C.POINTER(_Csr), f64p, f64p]),
            "orc_cg": (_Report, [C.POINTER(_Csr), f64p, f64p, C.c_double, C.c_int64, f64p]),
            "orc_minres": (_Report, [C.POINTER(_Csr), f64p, f64p, C.c_double, C.c_int64, f64p]),
            "orc_minres_cheb": (_Report, [C.POINTER(_Csr), C.POINTER(_Csr), f64p, C.c_int64, C.c_int,
                                          f64p, f64p, C.c_double, C.c_int64, f64p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if path != os.path.join(_HERE, "liboracle.so"):
            return L
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


# --------------------------------------------------------------------------
# CSR
# --------------------------------------------------------------------------

class Csr:
    """Owned copy of an oracle CSR: row_ptr, col (int64), val (float64)."""

    def __init__(self, shape, row_ptr, col, val):
        self.shape = (int(shape[0]), int(shape[1]))
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int64)
        self.val = np.ascontiguousarray(val, dtype=np.float64)

    @property
    def nnz(self):
        return int(self.col.size)

    @staticmethod
    def _take(ptr):
        s = ptr.contents
        n, nnz = s.nrows, s.nnz
        rp = np.ctypeslib.as_array(s.row_ptr, shape=(n + 1,)).copy() if n >= 0 else np.zeros(1, np.int64)
        col = np.ctypeslib.as_array(s.col, shape=(max(nnz, 1),))[:nnz].copy()
        val = np.ctypeslib.as_array(s.val, shape=(max(nnz, 1),))[:nnz].copy()
        out = Csr((s.nrows, s.ncols), rp, col, val)
        lib().orc_csr_free(ptr)
        return out

    def _struct(self):
        s = _Csr(self.shape[0], self.shape[1], self.nnz, _p(self.row_ptr, i64p),
                 _p(self.col, i64p), _p(self.val, f64p))
        s._keep = self
        return s

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.row_ptr), shape=self.shape)

    def todense(self):
        out = np.zeros(self.shape)
        for i in range(self.shape[0]):
            for p in range(self.row_ptr[i], self.row_ptr[i + 1]):
                out[i, self.col[p]] += self.val[p]
        return out

    def __add__(self, other):
        a, b = self._struct(), other._struct()
        return Csr._take(lib().orc_csr_add(C.byref(a), C.byref(b)))

    def __matmul__(self, other):
        if isinstance(other, np.ndarray):
            x = np.ascontiguousarray(other, dtype=np.float64)
            y = np.zeros(self.shape[0])
            a = self._struct()
            lib().orc_spmv(C.byref(a), _p(x, f64p), _p(y, f64p))
            return y
        a, b = self._struct(), other._struct()
        return Csr._take(lib().orc_csr_mul(C.byref(a), C.byref(b)))

    @property
    def T(self):
        a = self._struct()
        return Csr._take(lib().orc_csr_transpose(C.byref(a)))

    def select(self, row_keep=None, col_keep=None):
        a = self._struct()
        rk = None if row_keep is None else np.ascontiguousarray(row_keep, dtype=np.uint8)
        ck = None if col_keep is None else np.ascontiguousarray(col_keep, dtype=np.uint8)
        return Csr._take(lib().orc_csr_select(C.byref(a), None if rk is None else _p(rk, u8p),
                                              None if ck is None else _p(ck, u8p)))

    def diagonal(self):
        d = np.zeros(min(self.shape))
        for i in range(d.size):
            for p in range(self.row_ptr[i], self.row_ptr[i + 1]):
                if self.col[p] == i:
                    d[i] = self.val[p]
        return d

    @staticmethod
    def zeros(nrows, ncols):
        return Csr((nrows, ncols), np.zeros(nrows + 1, np.int64), np.zeros(0, np.int64), np.zeros(0))

    @staticmethod
    def from_dense_rows(dense):
        """Dense block kept structurally in full (used for harmonic borders)."""
        dense = np.asarray(dense, dtype=np.float64)
        n, m = dense.shape
        rp = np.arange(n + 1, dtype=np.int64) * m
        col = np.tile(np.arange(m, dtype=np.int64), n)
        return Csr((n, m), rp, col, dense.reshape(-1).copy())


def block(blocks, scales=None, row_sizes=None, col_sizes=None):
    """CooMatrix::block: a square grid of CSR blocks (None = empty)."""
    nb = len(blocks)
    flat = [blocks[i][j] for i in range(nb) for j in range(nb)]
    if row_sizes is None:
        row_sizes = [next(b.shape[0] for b in blocks[i] if b is not None) for i in range(nb)]
    if col_sizes is None:
        col_sizes = [next(blocks[i][j].shape[1] for i in range(nb) if blocks[i][j] is not None)
                     for j in range(nb)]
    structs = [None if b is None else b._struct() for b in flat]
    arr = (C.POINTER(_Csr) * (nb * nb))(*[C.cast(C.pointer(s), C.POINTER(_Csr)) if s is not None
                                          else C.POINTER(_Csr)() for s in structs])
    sc = np.ones(nb * nb) if scales is None else np.ascontiguousarray(np.asarray(scales, float).reshape(-1))
    rs = np.ascontiguousarray(row_sizes, dtype=np.int64)
    cs = np.ascontiguousarray(col_sizes, dtype=np.int64)
    res = lib().orc_csr_block(nb, arr, _p(sc, f64p), _p(rs, i64p), _p(cs, i64p))
    del structs
    return Csr._take(res)


# --------------------------------------------------------------------------
# meshes, topology, geometry
# --------------------------------------------------------------------------

def kuhn(dim, n):
    """CartesianGrid::new_unit(dim, n).triangulate_cells(): Kuhn cells in
    generation order and vertex coordinates [V, dim]."""
    L = lib()
    nc, nv = L.orc_kuhn_ncells(dim, n), L.orc_kuhn_nverts(dim, n)
    cells = np.zeros((nc, dim + 1), np.int64)
    coords = np.zeros((nv, dim), np.float64)
    L.orc_kuhn(dim, n, _p(cells, i64p), _p(coords, f64p))
    return cells, coords


def local_subsets(d, k):
    n = int(lib().orc_binom(d + 1, k + 1))
    out = np.zeros(n * (k + 1), dtype=np.int32)
    lib().orc_local_subsets(d, k, out.ctypes.data_as(C.POINTER(C.c_int)))
    return out.reshape(n, k + 1)


def form_shape(d, kind, grade):
    r, c = C.c_int(), C.c_int()
    lib().orc_form_shape(d, kind, grade, C.byref(r), C.byref(c))
    return r.value, c.value


class Complex:
    def __init__(self, dim, cells):
        cells = np.ascontiguousarray(cells, dtype=np.int64)
        h = lib().orc_complex_new(dim, cells.shape[0], _p(cells, i64p))
        if not h:
            raise ValueError("invalid cell list (dangling/duplicate vertices)")
        self._h = h
        self.dim = dim

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_complex_free(self._h)
            self._h = None

    def nsimplices(self, k):
        return int(lib().orc_nsimplices(self._h, k))

    def skeleton(self, k):
        out = np.zeros((self.nsimplices(k), k + 1), np.int64)
        lib().orc_skeleton(self._h, k, _p(out, i64p))
        return out

    def cells(self):
        return self.skeleton(self.dim)

    def cell_faces(self, k):
        nf = int(lib().orc_binom(self.dim + 1, k + 1))
        out = np.zeros((self.nsimplices(self.dim), nf), np.int64)
        lib().orc_cell_faces(self._h, k, _p(out, i64p))
        return out

    def boundary_flags(self, k):
        out = np.zeros(self.nsimplices(k), np.uint8)
        lib().orc_boundary_flags(self._h, k, _p(out, u8p))
        return out

    def coboundary(self, k):
        return Csr._take(lib().orc_coboundary(self._h, k))

    def orientation(self):
        out = np.zeros(self.nsimplices(self.dim), np.int8)
        ok = lib().orc_orientation(self._h, out.ctypes.data_as(C.POINTER(C.c_int8)))
        return out if ok else None

    def lengths_from_coords(self, coords, minkowski=False):
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        out = np.zeros(self.nsimplices(1))
        lib().orc_lengths_from_coords(self._h, _p(coords, f64p), int(minkowski), _p(out, f64p))
        return out


def cell_metric(d, s):
    s = np.ascontiguousarray(s, dtype=np.float64)
    g, gi, det = np.zeros(d * d), np.zeros(d * d), np.zeros(1)
    lib().orc_cell_metric(d, _p(s, f64p), _p(g, f64p), _p(gi, f64p), _p(det, f64p))
    return g.reshape(d, d), gi.reshape(d, d), float(det[0])


def element_matrix(d, kind, grade, s, L=None):
    r, c = form_shape(d, kind, grade)
    s = np.ascontiguousarray(s, dtype=np.float64)
    out = np.zeros(max(r * c, 1))
    (L or lib()).orc_element_matrix(d, kind, grade, _p(s, f64p), _p(out, f64p))
    return out[: r * c].reshape(r, c)


def unit_lengths(d):
    """SimplexLengthsSq::unit (regge/src/lengths/simplex.rs:81-88): edges from
    vertex 0 have squared length 1, the others 2."""
    edges = local_subsets(d, 1)
    return np.array([1.0 if e[0] == 0 else 2.0 for e in edges])


def element_matrices(cx, lengths, kind, grade, nthreads=0):
    r, c = form_shape(cx.dim, kind, grade)
    lengths = np.ascontiguousarray(lengths, dtype=np.float64)
    out = np.zeros((cx.nsimplices(cx.dim), r, c))
    lib().orc_element_matrices(cx._h, _p(lengths, f64p), kind, grade, _p(out, f64p), nthreads)
    return out


def assemble(cx, lengths, kind, grade, nthreads=0):
    lengths = np.ascontiguousarray(lengths, dtype=np.float64)
    return Csr._take(lib().orc_assemble(cx._h, _p(lengths, f64p), kind, grade, nthreads))


@dataclass
class Report:
    iters: int
    residual: float
    converged: bool


def _krylov(fn, a, b, inv_diag, rtol, max_iters):
    s = a._struct()
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(a.shape[0])
    idg = None if inv_diag is None else np.ascontiguousarray(inv_diag, dtype=np.float64)
    rep = fn(C.byref(s), None if idg is None else _p(idg, f64p), _p(b, f64p), rtol, max_iters, _p(x, f64p))
    return x, Report(int(rep.iters), float(rep.residual), bool(rep.converged))


def cg(a, b, inv_diag=None, rtol=1e-10, max_iters=10_000):
    return _krylov(lib().orc_cg, a, b, inv_diag, rtol, max_iters)


def minres(a, b, inv_diag=None, rtol=1e-10, max_iters=10_000):
    return _krylov(lib().orc_minres, a, b, inv_diag, rtol, max_iters)


# --------------------------------------------------------------------------
# the discrete Hilbert complex and the problems built on it
# --------------------------------------------------------------------------

class WhitneyComplex:
    """HilbertComplex over a Whitney complex (full or relative)."""

    def __init__(self, cx: Complex, lengths, relative=False, memo=False):
        self.cx = cx
        self.lengths = np.ascontiguousarray(lengths, dtype=np.float64)
        self.relative = relative
        # memo: keep each assembled (kind, grade) block for later requests
        self._memo = {} if memo else None
        self.dim = cx.dim
        self._interior = [None] * (cx.dim + 1)
        if relative:
            for k in range(cx.dim + 1):
                self._interior[k] = (cx.boundary_flags(k) == 0).astype(np.uint8)

    def in_range(self, k):
        return 0 <= k <= self.dim

    def ndofs(self, k):
        if not self.in_range(k):
            return 0
        if self.relative:
            return int(self._interior[k].sum())
        return self.cx.nsimplices(k)

    def interior(self, k):
        return self._interior[k]

    def _assemble(self, kind, grade):
        tg = grade - 1 if kind in (DIF_TEST, DIF_BOTH) else grade
        rg = grade - 1 if kind in (DIF_TRIAL, DIF_BOTH) else grade
        rows, cols = self.ndofs(tg), self.ndofs(rg)
        if rows == 0 or cols == 0:
            return Csr.zeros(rows, cols)
        if self._memo is not None and (kind, grade) in self._memo:
            return self._memo[(kind, grade)]
        a = assemble(self.cx, self.lengths, kind, grade)
        if self.relative:
            a = a.select(self._interior[tg], self._interior[rg])
        if self._memo is not None:
            self._memo[(kind, grade)] = a
        return a

    def mass(self, k):
        return self._assemble(MASS, k)

    def dif_trial(self, k):
        return self._assemble(DIF_TRIAL, k)

    def dif_test(self, k):
        return self._assemble(DIF_TEST, k)

    def dif_both(self, k):
        return self._assemble(DIF_BOTH, k)

    def hdif_gram(self, k):
        return self.mass(k) + self.dif_both(k + 1)

    def dif(self, k):
        if not (0 <= k < self.dim):
            return Csr.zeros(self.ndofs(k + 1), self.ndofs(k))
        d = self.cx.coboundary(k)
        if self.relative:
            d = d.select(self._interior[k + 1], self._interior[k])
        return d

    def inclusion_indices(self, k):
        if self.relative:
            return np.nonzero(self._interior[k])[0]
        return np.arange(self.ndofs(k))

    def harmonics(self, k, method="auto"):
        """An M_k-orthonormal harmonic basis (elliptic.rs:197-220).

        method "closed": the closed forms on a ball: b_0 = 1 on the full
        complex (the constant), b_n = 1 on the relative one (M_n^{-1} times the
        orientation, dirac.rs:553-568), zero otherwise. "exact": the projected
        integral cocycles of harmonic.rs:118-175 (``harmonics_exact``), for any
        topology. "auto": "closed" on connected complexes with Euler
        characteristic 1 (every such test mesh is a ball), else "exact"."""
        if method == "auto":
            method = "closed" if _looks_like_ball(self.cx) else "exact"
        if method == "exact":
            return harmonics_exact(self, k)[1]
        n = self.ndofs(k)
        if not self.relative and k == 0:
            h = np.ones(n)
        elif self.relative and k == self.dim:
            o = self.cx.orientation().astype(np.float64)
            m = self.mass(k).diagonal()
            h = o / m
        else:
            return np.zeros((n, 0))
        m = self.mass(k)
        h = h / math.sqrt(h @ (m @ h))
        return h.reshape(n, 1)


# ---------------------------------------------------------------------------
# Exact (co)homology and harmonic bases (simplicial/src/linalg/exact.rs,
# simplicial/src/topology/{homology,cohomology}.rs, formoniq/src/harmonic.rs)
# ---------------------------------------------------------------------------
def _looks_like_ball(cx):
    chi = sum((-1) ** k * cx.nsimplices(k) for k in range(cx.dim + 1))
    if chi != 1:
        return False
    # one component: vertices joined through the cells
    parent = list(range(cx.nsimplices(0)))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    for c in cx.cells():
        for v in c[1:]:
            ra, rb = find(int(c[0])), find(int(v))
            if ra != rb:
                parent[ra] = rb
    return len({find(v) for v in range(len(parent))}) == 1


def integral_coboundary(cx, k, relative=False):
    """D_k: C^k -> C^{k+1} as a dense integer matrix on the DOFs (relative:
    the interior simplices), the transpose of integral_boundary(k+1)
    (cohomology.rs:38-45); empty outside 0 <= k < dim."""
    def dofs(j):
        if j < 0 or j > cx.dim:
            return np.zeros(0, np.int64)
        if relative:
            return np.nonzero(cx.boundary_flags(j) == 0)[0]
        return np.arange(cx.nsimplices(j))
    rows, cols = dofs(k + 1), dofs(k)
    if k < 0 or k >= cx.dim or len(rows) == 0 or len(cols) == 0:
        return np.zeros((len(rows), len(cols)), np.int64)
    full = np.rint(cx.coboundary(k).todense()).astype(np.int64)
    return full[np.ix_(rows, cols)]


def rank_mod(a, p):
    """exact.rs:117-147: rows reduced in order against pivot rows keyed by
    their leading column, each pivot normalized to a leading 1 mod p."""
    pivots = {}
    rank = 0
    for row in np.asarray(a, dtype=np.int64) % p:
        row = row.copy()
        while True:
            nz = np.nonzero(row)[0]
            if len(nz) == 0:
                break
            lead = int(nz[0])
            if lead not in pivots:
                inv = pow(int(row[lead]), p - 2, p)
                pivots[lead] = (row * inv) % p
                rank += 1
                break
            row = (row - int(row[lead]) * pivots[lead]) % p
    return rank


def exact_rank(a):
    """exact.rs:106-115: the larger of the ranks over F_p for two primes."""
    a = np.asarray(a, dtype=np.int64)
    if a.size == 0:
        return 0
    return max(rank_mod(a, 1_000_000_007), rank_mod(a, 1_000_000_009))


def _rref(mat, ncols):
    """exact.rs:266-298 over Fractions: column by column, the first row at or
    below the current one with a nonzero entry is swapped up, normalized and
    eliminated from every other row."""
    from fractions import Fraction
    mat = [[Fraction(int(x)) for x in r] for r in mat]
    nrows = len(mat)
    pivot_cols = []
    row = 0
    for col in range(ncols):
        if row >= nrows:
            break
        sel = next((r for r in range(row, nrows) if mat[r][col] != 0), None)
        if sel is None:
            continue
        mat[row], mat[sel] = mat[sel], mat[row]
        inv = 1 / mat[row][col]
        mat[row] = [x * inv for x in mat[row]]
        for r in range(nrows):
            if r != row and mat[r][col] != 0:
                f = mat[r][col]
                mat[r] = [x - f * y for x, y in zip(mat[r], mat[row])]
        pivot_cols.append(col)
        row += 1
    return mat, pivot_cols


def exact_kernel(a):
    """exact.rs:149-175: one kernel vector per free column of the RREF, a 1
    there and minus the RREF entries at the pivot columns."""
    from fractions import Fraction
    a = np.asarray(a, dtype=np.int64)
    ncols = a.shape[1]
    mat, pivots = _rref(a.tolist(), ncols)
    is_pivot = set(pivots)
    out = []
    for free in range(ncols):
        if free in is_pivot:
            continue
        v = [Fraction(0)] * ncols
        v[free] = Fraction(1)
        for r, pc in enumerate(pivots):
            v[pc] = -mat[r][free]
        out.append(v)
    return out


def _primitive(v):
    """exact.rs:300-330: denominators cleared, gcd divided out, the first
    nonzero entry positive."""
    lcm = 1
    for x in v:
        lcm = lcm * x.denominator // math.gcd(lcm, x.denominator)
    ints = [int(x * lcm) for x in v]
    g = 0
    for x in ints:
        g = math.gcd(g, x)
    if g:
        ints = [x // g for x in ints]
    first = next((x for x in ints if x != 0), 0)
    if first < 0 and MUTATION != 52:
        ints = [-x for x in ints]
    return ints


def quotient_generators(a, b):
    """exact.rs:178-264: representatives of ker A / im B. An echelon span is
    seeded with B's columns; each RREF kernel vector of A, in free-column
    order, is reduced against it and kept (primitive) when it enlarges it."""
    from fractions import Fraction
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    assert a.shape[1] == b.shape[0]
    pivots = {}

    def insert(vec):
        red = list(vec)
        while True:
            lead = next((i for i, x in enumerate(red) if x != 0), None)
            if lead is None:
                return False
            if lead not in pivots:
                inv = 1 / red[lead]
                pivots[lead] = [x * inv for x in red]
                return True
            f = red[lead]
            red = [x - f * y for x, y in zip(red, pivots[lead])]

    for j in range(b.shape[1]):
        insert([Fraction(int(x)) for x in b[:, j]])
    out = []
    for v in exact_kernel(a):
        grew = insert(v)
        if grew != (MUTATION == 51):
            out.append(_primitive(v))
    return out


def betti_number(cx, k, relative=False):
    """homology.rs:47-52 / 81-98: n_k - rank d_k - rank d_{k+1}, d_k = D_{k-1}^T."""
    lo, hi = integral_coboundary(cx, k - 1, relative), integral_coboundary(cx, k, relative)
    return lo.shape[0] - exact_rank(lo) - exact_rank(hi)


def cohomology_generators(cx, k, relative=False):
    """cohomology.rs:55-90 on the DOFs: ker D_k / im D_{k-1}."""
    return quotient_generators(integral_coboundary(cx, k, relative), integral_coboundary(cx, k - 1, relative))


def homology_generators(cx, k, relative=False):
    """homology.rs:116-153 on the DOFs: ker d_k / im d_{k+1}."""
    return quotient_generators(integral_coboundary(cx, k - 1, relative).T,
                               integral_coboundary(cx, k, relative).T)


def harmonics_exact(c: WhitneyComplex, k):
    """harmonic.rs:118-175: (integral, orthonormal), each ndofs x b_k. Every
    integral cocycle z is projected to h = z - D p, D = dif(k-1), p minimizing
    |z - D p|_M (the reference's CG on D^T M D p = D^T M z; any solution gives
    the same h), then integral = H (P^T)^-1 with P_ij = <z_i, c_j> against the
    integral cycles, and orthonormal = integral L^-T for G = L L^T the Gram
    matrix in M. Raises ValueError where G is not positive definite."""
    z = cohomology_generators(c.cx, k, c.relative)
    cyc = homology_generators(c.cx, k, c.relative)
    n = c.ndofs(k)
    if not z:
        return np.zeros((n, 0)), np.zeros((n, 0))
    M = c.mass(k).todense()
    Z = np.array(z, dtype=np.float64).T
    D = c.dif(k - 1).todense() if k >= 1 else np.zeros((n, 0))
    if D.shape[1]:
        Lm = np.linalg.cholesky(M)
        p, *_ = np.linalg.lstsq(Lm.T @ D, Lm.T @ Z, rcond=None)
        H = Z + D @ p if MUTATION == 54 else Z - D @ p
    else:
        H = Z
    P = np.array([[sum(int(x) * int(y) for x, y in zip(zi, cj)) for cj in cyc] for zi in z], dtype=np.float64)
    integral = H if MUTATION == 53 else H @ np.linalg.inv(P.T)
    G = integral.T @ M @ integral
    L = np.linalg.cholesky(G)
    ortho = np.linalg.solve(L, integral.T).T
    return integral, ortho


def hodge_blocks(c: WhitneyComplex, k):
    return dict(n_sigma=c.ndofs(k - 1), n_u=c.ndofs(k), mass_sigma=c.mass(k - 1),
                mass_u=c.mass(k), dif_test=c.dif_test(k), dif_both=c.dif_both(k + 1))


def assemble_mixed_kkt(c: WhitneyComplex, source, k, harmonics):
    """elliptic.rs:54-114: [[-M_{k-1}, B, 0], [B^T, K, M H], [0, H^T M, 0]],
    rhs [0; E^T f; 0]."""
    b = hodge_blocks(c, k)
    ns, nu, nh = b["n_sigma"], b["n_u"], harmonics.shape[1]
    mh = b["mass_u"].to_scipy() @ harmonics if nh else np.zeros((nu, 0))
    blocks = [[b["mass_sigma"], b["dif_test"], None],
              [b["dif_test"].T, b["dif_both"], Csr.from_dense_rows(mh) if nh else None],
              [None, Csr.from_dense_rows(mh.T) if nh else None, None]]
    # mixed_hodge_laplacian has -dif_test in the sigma row; the sign flip of
    # the sigma rows turns [M, -B] into [-M, B].
    scales = [[-1.0, 1.0, 1.0], [1.0, 1.0, 1.0], [1.0, 1.0, 1.0]]
    if MUTATION == 21:
        scales[0][1] = -1.0
    if MUTATION == 22:
        scales[0][0] = -2.0
    if MUTATION == 23:
        scales[1][1] = -1.0
    kkt = block(blocks, scales, [ns, nu, nh], [ns, nu, nh])
    src = np.asarray(source, dtype=np.float64)[c.inclusion_indices(k)]
    rhs = np.concatenate([np.zeros(ns), src, np.zeros(nh)])
    return kkt, rhs, ns, nu


def mixed_jacobi(c: WhitneyComplex, k, nh):
    """Inverse diagonal of the block preconditioner diag(hdif_gram(k-1),
    hdif_gram(k), I) (elliptic.rs:29-47), Jacobi-approximated per block."""
    parts = []
    if k > 0:
        parts.append(1.0 / c.hdif_gram(k - 1).diagonal())
    parts.append(1.0 / c.hdif_gram(k).diagonal())
    parts.append(np.ones(nh))
    return np.concatenate(parts)


def solve_source(c: WhitneyComplex, source, k, precond="jacobi", rtol=1e-10, max_iters=100_000):
    """elliptic.rs:132-182 with the block preconditioner applied by Jacobi
    ("jacobi") or by sparse LU of each block ("direct", the reference's
    DirectInverse with SuperLU standing in for faer). Returns sigma, u, p (ambient
    cochains) and the MINRES report."""
    h = c.harmonics(k)
    kkt, rhs, ns, nu = assemble_mixed_kkt(c, source, k, h)
    if precond == "jacobi":
        x, rep = minres(kkt, rhs, mixed_jacobi(c, k, h.shape[1]), rtol, max_iters)
    elif precond == "direct":
        x, rep = _minres_direct_blocks(c, k, kkt, rhs, h.shape[1], rtol, max_iters)
    else:
        raise ValueError(precond)
    sig = np.zeros(c.cx.nsimplices(k - 1) if k > 0 else 0)
    if k > 0:
        sig[c.inclusion_indices(k - 1)] = x[:ns]
    u = np.zeros(c.cx.nsimplices(k))
    u[c.inclusion_indices(k)] = x[ns:ns + nu]
    return sig, u, x[ns + nu:], rep


def _minres_direct_blocks(c, k, kkt, rhs, nh, rtol, max_iters):
    """MINRES (krylov.rs:113-211) in Python with the block-direct
    preconditioner; SuperLU stands in for faer's Cholesky."""
    import scipy.sparse.linalg as spla
    facs = []
    if k > 0:
        facs.append(spla.splu(c.hdif_gram(k - 1).to_scipy().tocsc()))
    facs.append(spla.splu(c.hdif_gram(k).to_scipy().tocsc()))
    sizes = [f.shape[0] for f in facs] + [nh]

    def prec(r):
        out, off = np.empty_like(r), 0
        for i, n in enumerate(sizes):
            out[off:off + n] = facs[i].solve(r[off:off + n]) if i < len(facs) else r[off:off + n]
            off += n
        return out

    A = kkt.to_scipy()
    return minres_py(lambda v: A @ v, prec, rhs, rtol, max_iters)


CHEB_POWER_ITERS = 30
CHEB_LMAX_MARGIN = 1.05
CHEB_GERSHGORIN_TIGHT = 1.1


def cheb_interval_top(lmax, gershgorin):
    """The Chebyshev interval's upper end: the Gershgorin bound of D^-1 P
    (max_i sum_j |P_ij| / P_ii, never below lambda_max) where it lies within
    10% of the power estimate, else 1.05 times the estimate."""
    if 0.0 < gershgorin <= CHEB_GERSHGORIN_TIGHT * lmax:
        return gershgorin
    return CHEB_LMAX_MARGIN * lmax


def cheb_interval_fraction(m):
    """Lower end of the Chebyshev interval as a fraction of its upper end."""
    return (1e-2 if m <= 4 else 3e-3 if m <= 8 else 1.5e-3 if m <= 12 else 1e-3 if m <= 32
            else 5e-4 if m <= 48 else 2.5e-4 if m <= 64 else 1.2e-4 if m <= 96 else 6e-5)


def _xorshift_start(n):
    """Power-iteration start vector: 0.5 + U[0,1) from xorshift64 (13, 7, 17)."""
    out = np.empty(n)
    s = 0x2545F4914F6CDD1D
    mask = (1 << 64) - 1
    for i in range(n):
        s ^= (s << 13) & mask
        s ^= s >> 7
        s ^= (s << 17) & mask
        out[i] = 0.5 + (s >> 11) * (1.0 / 9007199254740992.0)
    return out


def chebyshev_coefficients(lmax, m, top=None):
    """(1/theta, [(rho_{k+1} rho_k, 2 rho_{k+1}/delta)] for k < m-1) of the
    Chebyshev iteration on [f b, b], b = top, or 1.05 lmax without it (Saad,
    Iterative Methods for Sparse Linear Systems, 2nd ed., Algorithm 12.1)."""
    b = CHEB_LMAX_MARGIN * lmax if top is None else top
    a = b * cheb_interval_fraction(m)
    theta, delta = 0.5 * (b + a), 0.5 * (b - a)
    sigma = theta / delta
    rho = 1.0 / sigma
    steps = []
    for _ in range(m - 1):
        rn = 1.0 / (2.0 * sigma - rho)
        steps.append((rn * rho, 2.0 * rn / delta))
        rho = rn
    return 1.0 / theta, steps


def chebyshev_apply(P, dinv, r, c0, steps):
    """z = p(D^-1 P) D^-1 r by the Chebyshev recurrence: m directions summed,
    so 1 - lambda p(lambda) = T_m((theta - lambda)/delta) / T_m(theta/delta)."""
    res = dinv * r
    d = c0 * res
    z = np.zeros_like(r)
    for a, b in steps:
        res = res - dinv * (P @ d)
        z = z + d
        d = a * d + b * res
    return z + d


def chebyshev_apply_three_term(P, dinv, r, c0, steps):
    """The same z by the form the mixed GPU path evaluates: on the similar
    matrix S = D^-1/2 P D^-1/2 (unit diagonal) with f = D^-1/2 r, x_0 = 0,
    x_1 = c0 f, x_{k+2} = x_{k+1} + a_k (x_{k+1} - x_k) + b_k (f - S x_{k+1})
    (d_{k+1} = x_{k+2} - x_{k+1}, r_{k+1} = f - S x_{k+1}), z = D^-1/2 x_m,
    since p(D^-1 P) D^-1 = D^-1/2 p(S) D^-1/2. The diagonal of S is taken
    as exactly 1 and only P's off-diagonal part is applied."""
    import scipy.sparse as sp
    sd = np.sqrt(np.abs(dinv))
    S = sp.diags(sd) @ P @ sp.diags(sd)
    S = S - sp.diags(S.diagonal())
    f = sd * r
    x0, x1 = np.zeros_like(f), c0 * f
    for a, b in steps:
        x0, x1 = x1, x1 + a * (x1 - x0) + b * (f - (x1 + S @ x1))
    return sd * x1


def chebyshev_jacobi(c: WhitneyComplex, k, nh, m):
    """Build-defined stand-in for the reference's direct block preconditioner
    diag(hdif_gram(k-1)^-1, hdif_gram(k)^-1, I) (elliptic.rs:29-47): per
    block, a degree-(m-1) Chebyshev polynomial in D^-1 hdif_gram with D its
    diagonal, lambda_max from 30 power iterations and the interval's top
    from cheb_interval_top. Returns prec(r)."""
    import scipy.sparse as sp
    blocks = ([c.hdif_gram(k - 1).to_scipy()] if k > 0 else []) + [c.hdif_gram(k).to_scipy()]
    P = sp.block_diag(blocks, format="csr")
    n = P.shape[0]
    dinv = 1.0 / P.diagonal()
    split = blocks[0].shape[0] if k > 0 else 0
    bounds = [(0, split), (split, n)]
    x = _xorshift_start(n)
    lam = [1.0, 1.0]
    for _ in range(CHEB_POWER_ITERS):
        y = dinv * (P @ x)
        for b, (lo, hi) in enumerate(bounds):
            if hi <= lo:
                continue
            ny, nx = np.linalg.norm(y[lo:hi]), np.linalg.norm(x[lo:hi])
            if nx > 0:
                lam[b] = ny / nx
            if ny > 0:
                x[lo:hi] = y[lo:hi] / ny
    rows = np.abs(P).sum(axis=1).A1 * dinv
    gersh = [rows[lo:hi].max() if hi > lo else 0.0 for lo, hi in bounds]
    coefs = [chebyshev_coefficients(lam[b], m, cheb_interval_top(lam[b], gersh[b])) for b in range(2)]
    Pb = [blocks[0] if k > 0 else None, blocks[-1]]

    def prec(r):
        out = np.empty_like(r)
        for b, (lo, hi) in enumerate(bounds):
            if hi > lo:
                out[lo:hi] = chebyshev_apply(Pb[b], dinv[lo:hi], r[lo:hi], *coefs[b])
        out[n:] = r[n:]
        return out

    prec.lambda_max = lam
    prec.P = Csr(P.shape, P.indptr.astype(np.int64), P.indices.astype(np.int64), P.data)
    prec.dinv = dinv
    prec.split = split
    prec.m = m
    flat = [coefs[0][0], coefs[1][0]]
    for k in range(m - 1):
        for b in range(2):
            flat += list(coefs[b][1][k])
    prec.coef = np.array(flat, dtype=np.float64)
    return prec


def minres_cheb(a, b, prec, rtol=1e-10, max_iters=10_000):
    """The C oracle's MINRES with the Chebyshev block preconditioner built by
    chebyshev_jacobi (the same operator as its Python closure)."""
    s, ps = a._struct(), prec.P._struct()
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(a.shape[0])
    rep = lib().orc_minres_cheb(C.byref(s), C.byref(ps), _p(prec.dinv, f64p), prec.split, prec.m,
                                _p(prec.coef, f64p), _p(b, f64p), rtol, max_iters, _p(x, f64p))
    return x, Report(int(rep.iters), float(rep.residual), bool(rep.converged))


def minres_py(op, prec, b, rtol=1e-10, max_iters=10_000):
    """krylov.rs:113-211 with callables (for non-diagonal preconditioners)."""
    r1 = b.copy()
    y = prec(r1)
    beta1_sq = r1 @ y
    x = np.zeros_like(b)
    if beta1_sq <= 0:
        return x, Report(0, 0.0, True)
    beta1 = math.sqrt(beta1_sq)
    oldb, beta, dbar, epsln, phibar, cs, sn = 0.0, beta1, 0.0, 0.0, beta1, -1.0, 0.0
    w, w2, r2 = np.zeros_like(b), np.zeros_like(b), r1.copy()
    residual, iters = 1.0, 0
    while iters < max_iters:
        iters += 1
        v = y * (1.0 / beta)
        yn = op(v)
        if iters >= 2:
            yn = yn + (-beta / oldb) * r1
        alfa = v @ yn
        yn = yn + (-alfa / beta) * r2
        r1, r2 = r2, yn
        y = prec(r2)
        oldb = beta
        beta = math.sqrt(max(r2 @ y, 0.0))
        oldeps = epsln
        delta = cs * dbar + sn * alfa
        gbar = sn * dbar - cs * alfa
        epsln = sn * beta
        dbar = -cs * beta
        gamma = max(math.sqrt(gbar * gbar + beta * beta), 2.220446049250313e-16)
        cs, sn = gbar / gamma, beta / gamma
        phi = cs * phibar
        phibar *= sn
        wn = (v - oldeps * w2 - delta * w) * (1.0 / gamma)
        w2, w = w, wn
        x = x + phi * wn
        residual = phibar / beta1
        if residual <= rtol:
            return x, Report(iters, residual, True)
    return x, Report(iters, residual, False)


def hodge_dirac(c: WhitneyComplex, sign):
    """HodgeDirac::assemble_signed (dirac.rs:213-258): offsets, mass_block and
    op with U_k = M_k D_{k-1} below the diagonal and sign * U_k^T above."""
    n = c.dim
    masses = [c.mass(k) for k in range(n + 1)]
    difs = [c.dif(k) for k in range(n)]
    offsets = np.cumsum([0] + [m.shape[0] for m in masses])
    nb = n + 1
    mb = [[masses[i] if i == j else None for j in range(nb)] for i in range(nb)]
    sizes = [m.shape[0] for m in masses]
    mass_block = block(mb, None, sizes, sizes)
    # op: emit (r,c,v) and (c,r,sign v) from the same triplets, then compress.
    import scipy.sparse as sp
    rows, cols, vals = [], [], []
    for k in range(1, n + 1):
        u = masses[k] @ difs[k - 1]
        if MUTATION == 11:
            u = difs[k - 1] @ masses[k - 1]
        if MUTATION == 12:
            u = Csr(u.shape, u.row_ptr, u.col, 2.0 * u.val)
        coo_r = np.repeat(np.arange(u.shape[0]), np.diff(u.row_ptr))
        r0, c0 = offsets[k], offsets[k - 1]
        rows += [r0 + coo_r, c0 + u.col]
        cols += [c0 + u.col, r0 + coo_r]
        vals += [-u.val if MUTATION == 13 else u.val, sign * u.val]
    total = int(offsets[-1])
    if rows:
        r = np.concatenate(rows)
        cc = np.concatenate(cols)
        v = np.concatenate(vals)
    else:
        r = cc = np.zeros(0, np.int64)
        v = np.zeros(0)
    # No duplicate (r, c) pairs across the emitted blocks, so compression is a
    # pure reorder and scipy's conversion changes no value.
    op = sp.csr_matrix((v, (r, cc)), shape=(total, total))
    op.sort_indices()
    opc = Csr((total, total), op.indptr.astype(np.int64), op.indices.astype(np.int64), op.data)
    return offsets, mass_block, opc


def grade_parity_coloring(offsets):
    """HodgeDirac::grade_parity_coloring (dirac.rs:330-337): True on the DOFs
    of odd grades."""
    color = np.zeros(int(offsets[-1]), bool)
    for k in range(1, len(offsets) - 1, 2):
        color[offsets[k]:offsets[k + 1]] = True
    return color


class Leapfrog:
    """Leapfrog::{new, step, conserved_energy} (time.rs:292-420): the
    color-block masses and couplings of mass and op, the color-block masses
    factored once (scipy SuperLU standing in for the reference's Cholesky)."""

    def __init__(self, mass, op, color, dt):
        import scipy.sparse.linalg as spla
        m, a = mass.to_scipy().tocsr(), op.to_scipy().tocsr()
        self.i0, self.i1 = np.flatnonzero(~color), np.flatnonzero(color)
        mc, ac = m.tocoo(), a.tocoo()
        # the split is valid only if M keeps and A swaps the colors (time.rs:330-343)
        assert not np.any((color[mc.row] != color[mc.col]) & (mc.data != 0.0)), "mass couples the colors"
        assert not np.any((color[ac.row] == color[ac.col]) & (ac.data != 0.0)), "op couples within a color"
        self.m0 = m[self.i0][:, self.i0].tocsc()
        self.m1 = m[self.i1][:, self.i1].tocsc()
        self.a01 = a[self.i0][:, self.i1].tocsr()
        self.a10 = a[self.i1][:, self.i0].tocsr()
        self.lu0, self.lu1 = spla.splu(self.m0), spla.splu(self.m1)
        self.dt = dt
        self.n = m.shape[0]

    def step(self, y):
        q, p = y[self.i0].copy(), y[self.i1].copy()
        p += 0.5 * self.dt * self.lu1.solve(self.a10 @ q)
        q += self.dt * self.lu0.solve(self.a01 @ p)
        p += 0.5 * self.dt * self.lu1.solve(self.a10 @ q)
        out = np.zeros(self.n)
        out[self.i0], out[self.i1] = q, p
        return out

    def conserved_energy(self, y):
        q, p = y[self.i0], y[self.i1]
        a10q = self.a10 @ q
        return (0.5 * q @ (self.m0 @ q) + 0.5 * p @ (self.m1 @ p)
                - self.dt ** 2 / 8.0 * a10q @ self.lu1.solve(a10q))


def dirac_leapfrog(c: WhitneyComplex, dt, nsteps, y0):
    """solve_dirac_leapfrog's loop (dirac.rs:415-434) on the flat field of the
    complex: the final state and the invariant of every state."""
    offsets, mass, op = hodge_dirac(c, -1.0)
    lf = Leapfrog(mass, op, grade_parity_coloring(offsets), dt)
    y = np.array(y0, dtype=np.float64)
    energy = [lf.conserved_energy(y)]
    for _ in range(nsteps):
        y = lf.step(y)
        energy.append(lf.conserved_energy(y))
    return y, np.array(energy)


def gauss_legendre(s):
    """Tableau::gauss_legendre (time.rs:59-67) for s = 1, 2: (A, b, c) of the
    Gauss collocation methods, from the textbook nodes 1/2 -+ sqrt(3)/6."""
    if s == 1:
        return np.array([[0.5]]), np.array([1.0]), np.array([0.5])
    r = np.sqrt(3.0) / 6.0
    return (np.array([[0.25, 0.25 - r], [0.25 + r, 0.25]]), np.array([0.5, 0.5]),
            np.array([0.5 - r, 0.5 + r]))


class LinearIrk:
    """LinearIrk::{new, step} (time.rs:191-274) with zero forcing: the sparse
    stage matrix I_s (x) M - dt (A_tab (x) A) factored once (SuperLU standing in
    for faer's LU); a step solves it for the stage derivatives against
    1_s (x) A y0 and returns y0 + dt sum_i b_i K_i."""

    def __init__(self, stages, mass, op, dt):
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        self.a, self.b, _ = gauss_legendre(stages)
        m, a = mass.to_scipy().tocsr(), op.to_scipy().tocsr()
        s = self.a.shape[0]
        stage = sp.kron(sp.identity(s), m) - dt * sp.kron(sp.csr_matrix(self.a), a)
        self.lu = spla.splu(stage.tocsc())
        self.op, self.m, self.dt, self.s, self.n = a, m, dt, s, m.shape[0]

    def step(self, y0):
        k = self.lu.solve(np.tile(self.op @ y0, self.s))
        return y0 + self.dt * sum(self.b[i] * k[i * self.n:(i + 1) * self.n] for i in range(self.s))

    def energy(self, y):
        return 0.5 * y @ (self.m @ y)


def dirac_irk(c: WhitneyComplex, dt, nsteps, y0, stages=2):
    """solve_dirac's loop (dirac.rs:383-403) on the flat field of the complex:
    the final state and the energy 1/2 y^T M y of every state."""
    offsets, mass, op = hodge_dirac(c, -1.0)
    irk = LinearIrk(stages, mass, op, dt)
    y = np.array(y0, dtype=np.float64)
    energy = [irk.energy(y)]
    for _ in range(nsteps):
        y = irk.step(y)
        energy.append(irk.energy(y))
    return y, np.array(energy)


# ---------------------------------------------------------------------------
# Mesh-independent block preconditioner: geometric multigrid on Kuhn grids
# inside the Hiptmair-Xu auxiliary-space preconditioner
#   VCycle::cycle           iterative/src/multigrid.rs (smooth, residual, restrict,
#                           coarse correction, prolong, smooth; direct coarsest)
#   Jacobi::smoother        iterative/src/precond.rs:89-94 (omega = 2/3)
#   stationary::sweeps      iterative/src/stationary.rs:61-74
#   AuxiliarySpace          iterative/src/aux_space.rs (B = S + sum Pi B_i Pi^T)
#   hx_preconditioner       formoniq/src/hx.rs:244-275 (recursing in grade)
#   vector_nodal_prolongation  formoniq/src/hx.rs:86-112
#   RefinementTower         formoniq/src/multigrid.rs:55-138 (the reference refines
#                           a base mesh; on a Kuhn grid of n cells per axis the
#                           tower is recovered by coarsening to ceil(n/2), with
#                           Galerkin coarse operators P^T A P)
# ---------------------------------------------------------------------------
MG_COARSE_MAX = 200  # a level with at most this many unknowns is solved directly
MG_OMEGA = 2.0 / 3.0


def _grid_coords(d, g):
    ids = np.arange((g + 1) ** d)
    return ids, np.stack([(ids // (g + 1) ** a) % (g + 1) for a in range(d)], axis=1)


def kuhn_nodal_prolongation(d, g, interior=False):
    """Nodal prolongation from the Kuhn grid of ceil(g/2) cells per axis to
    that of g (vertex c at sum_a c_a (g+1)^a): fine vertex 2w + delta, delta in
    {0,1}^d, is the midpoint of the coarse Kuhn edge (w, w + delta) and takes
    1 (delta = 0) or 1/2, 1/2 of its ends. `interior`: rows the fine interior
    vertices, columns the coarse vertices with 0 < w and 2 w <= g - 1 (the
    interior of the coarse grid), entries elsewhere dropped."""
    import scipy.sparse as sp
    gc = (g + 1) // 2
    ids, c = _grid_coords(d, g)
    delta = c % 2
    lo = (c - delta) // 2
    hi = lo + delta
    cid = lambda q: sum(q[:, a] * (gc + 1) ** a for a in range(d))
    z = delta.sum(1) == 0
    rows = np.concatenate([ids[z], ids[~z], ids[~z]])
    cols = np.concatenate([cid(lo)[z], cid(lo)[~z], cid(hi)[~z]])
    w_lo, w_hi = (0.75, 0.25) if MUTATION == 31 else (0.5, 0.5)
    vals = np.concatenate([np.ones(z.sum()), np.full((~z).sum(), w_lo), np.full((~z).sum(), w_hi)])
    P = sp.csr_matrix((vals, (rows, cols)), shape=((g + 1) ** d, (gc + 1) ** d))
    if interior:
        fin = np.all((c > 0) & (c < g), axis=1)
        _, w = _grid_coords(d, gc)
        cin = np.all((w > 0) & (2 * w <= g - 1), axis=1)
        P = P[fin][:, cin]
    return P.tocsr()


class NodalVCycle:
    """One V-cycle of the nodal problem A0 (hdif_gram(0) of the full or the
    relative complex on the Kuhn grid of n cells per axis): levels while a
    level has more than MG_COARSE_MAX unknowns, Galerkin coarse operators,
    weighted Jacobi smoothing with `sweeps` sweeps before and after the coarse
    correction, the coarsest level inverted directly."""

    def __init__(self, A0, d, n, relative, sweeps=2):
        self.sweeps = sweeps
        self.levels = []
        A = A0.tocsr()
        g = n
        while A.shape[0] > MG_COARSE_MAX and g >= 2:
            P = kuhn_nodal_prolongation(d, g, relative)
            self.levels.append((A, MG_OMEGA / A.diagonal(), P))
            A = (P.T @ (A @ P)).tocsr()
            g = (g + 1) // 2
        self.coarse = np.linalg.inv(A.toarray())

    def _cycle(self, i, r):
        if i == len(self.levels):
            return self.coarse @ r
        A, dinv, P = self.levels[i]
        x = np.zeros_like(r)
        for _ in range(self.sweeps):
            x = x + dinv * (r - A @ x)
        x = x + P @ self._cycle(i + 1, P.T @ (r - A @ x))
        for _ in range(self.sweeps if MUTATION != 32 else 0):
            x = x + dinv * (r - A @ x)
        return x

    def apply(self, r):
        return self._cycle(0, r)


def vector_nodal_prolongation(cx: Complex, coords, k):
    """hx.rs:86-112: the de Rham map of phi_a e_I, column I * N_0 + a (I the
    colex k-subsets of the d axes): on a k-simplex sigma containing a,
    (1/(k+1)!) <e_I, tau_sigma>, tau_sigma the wedge of sigma's edge vectors
    from its first vertex. Identity at k = 0."""
    import itertools
    import scipy.sparse as sp
    d = cx.dim
    n0 = cx.nsimplices(0)
    if k == 0:
        return sp.identity(n0, format="csr")
    simp = cx.skeleton(k)
    E = coords[simp[:, 1:]] - coords[simp[:, :1]]
    factor = 1.0 / math.factorial(k + (0 if MUTATION == 33 else 1))
    rows, cols, vals = [], [], []
    for bi, I in enumerate(itertools.combinations(range(d), k)):
        comp = np.linalg.det(E[:, :, list(I)])
        for p in range(k + 1):
            rows.append(np.arange(len(simp)))
            cols.append(bi * n0 + simp[:, p])
            vals.append(factor * comp)
    nb = math.comb(d, k)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(len(simp), nb * n0))


class HxPreconditioner:
    """hx_preconditioner (hx.rs:244-275) with multigrid blocks (hx.rs:302-313):
    B_k = S_k + D_{k-1} B_{k-1} D_{k-1}^T + Pi_k (V (x) I) Pi_k^T, B_0 = S_0 + V,
    S_j = 2/3 diag(hdif_gram(j))^-1, V the nodal V-cycle."""

    def __init__(self, c: WhitneyComplex, coords, k, n, sweeps=2, nodal=None):
        import scipy.sparse as sp
        self.S = MG_OMEGA / c.hdif_gram(k).diagonal()
        self.nodal = nodal or NodalVCycle(c.hdif_gram(0).to_scipy(), c.dim, n, c.relative, sweeps)
        self.grad = None
        if k > 0 and c.ndofs(k - 1) > 0:
            self.D = c.dif(k - 1).to_scipy().tocsr()
            self.grad = HxPreconditioner(c, coords, k - 1, n, sweeps, self.nodal)
        Pi = vector_nodal_prolongation(c.cx, coords, k)
        n0 = c.cx.nsimplices(0)
        self.Pi = []
        for b in range(Pi.shape[1] // n0):
            blk = Pi[:, b * n0:(b + 1) * n0]
            if c.relative:
                blk = blk[c.interior(k) == 1][:, c.interior(0) == 1]
            self.Pi.append(sp.csr_matrix(blk))

    def apply(self, r):
        z = self.S * r
        if self.grad is not None:
            z = z + self.D @ self.grad.apply(self.D.T @ r)
        for P in self.Pi:
            z = z + P @ self.nodal.apply(P.T @ r)
        return z


HX_CHEB_FRACTION = 0.25  # lower end of the interval, a fraction of its upper end


HX_LANCZOS_STEPS = 6
# a block stops once r.Br < this fraction of its start (FP32 path, FP64 path)
HX_LANCZOS_STOP = {True: 1e-10, False: 1e-20}


def tridiagonal_lmax(diag, off):
    """Largest eigenvalue of the symmetric tridiagonal matrix (diag, off) by
    bisection on Sturm counts inside the Gershgorin bound (60 halvings)."""
    n = len(diag)
    lo = min(diag[i] - (abs(off[i - 1]) if i else 0) - (abs(off[i]) if i + 1 < n else 0) for i in range(n))
    hi = max(diag[i] + (abs(off[i - 1]) if i else 0) + (abs(off[i]) if i + 1 < n else 0) for i in range(n))
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        count, q = 0, 1.0
        for i in range(n):  # eigenvalues below mid = negative pivots of T - mid I
            q = diag[i] - mid - (off[i - 1] ** 2 / q if i else 0.0)
            if q == 0.0:
                q = -1e-300
            count += q < 0
        if count == n:
            hi = mid
        else:
            lo = mid
    return hi


def single_symmetric(A):
    """The FP32 path's copy of a preconditioner block for its Chebyshev
    steps: (A + A^T)/2 rounded to FP32, returned in FP64."""
    S = (0.5 * (A + A.T)).tocsr()
    S.data = S.data.astype(np.float32).astype(np.float64)
    return S


def hx_blocks(c: WhitneyComplex, coords, k, n, sweeps=2):
    """The HX preconditioners (B_{k-1} or None, B_k) of the two KKT blocks,
    sharing one nodal V-cycle."""
    hu = HxPreconditioner(c, coords, k, n, sweeps)
    hs = HxPreconditioner(c, coords, k - 1, n, sweeps, hu.nodal) if c.ndofs(k - 1) else None
    return hs, hu


def hx_block_lambdas(c: WhitneyComplex, coords, k, n, steps=HX_LANCZOS_STEPS, sweeps=2, blocks=None,
                     single=False):
    """lambda_max of B_j hdif_gram(j) per block (sigma, u): `steps` steps of
    CG on hdif_gram(j) preconditioned by B_j from the xorshift start (fewer
    once r.Br < HX_LANCZOS_STOP of its start or p.Ap stops being positive;
    `single`: the FP32 path's estimate, on its FP32 copy of hdif_gram),
    and the largest eigenvalue of the Lanczos tridiagonal their coefficients
    make (T_jj = 1/alpha_j + beta_{j-1}/alpha_{j-1},
    T_{j,j+1} = sqrt(beta_j)/alpha_j)."""
    ns, nu = c.ndofs(k - 1), c.ndofs(k)
    hs, hu = blocks or hx_blocks(c, coords, k, n, sweeps)
    conv = single_symmetric if single else (lambda A: A)
    blocks = [(hs, conv(c.hdif_gram(k - 1).to_scipy().tocsr()) if ns else None, 0, ns),
              (hu, conv(c.hdif_gram(k).to_scipy().tocsr()), ns, ns + nu)]
    v = _xorshift_start(ns + nu)
    lam = [0.0, 0.0]
    for b, (B, A, lo, hi) in enumerate(blocks):
        if hi <= lo:
            continue
        r = v[lo:hi].copy()
        z = B.apply(r)
        p = z.copy()
        rho0 = rho = r @ z
        al, be = [], []
        for _ in range(steps):
            q = A @ p
            pw = p @ q
            if not pw > 0:  # the block's CG has run into roundoff
                break
            a = rho / pw
            al.append(a)
            r = r - a * q
            z = B.apply(r)
            rn = r @ z
            if not rn > HX_LANCZOS_STOP[single] * rho0:  # converged
                break
            be.append(rn / rho)
            rho = rn
            p = z + be[-1] * p
        m = len(al)
        diag = [1.0 / al[j] + (be[j - 1] / al[j - 1] if j else 0.0) for j in range(m)]
        off = [math.sqrt(be[j]) / al[j] for j in range(m - 1)]
        lam[b] = tridiagonal_lmax(diag, off)
    return lam


def hx_block_preconditioner(c: WhitneyComplex, coords, k, n, nh, sweeps=2, cheb_m=1, lam=None,
                            frac=HX_CHEB_FRACTION, blocks=None, single=False):
    """The KKT block preconditioner diag(B_{k-1}, B_k, I) (elliptic.rs:29-47)
    with HX blocks; a callable r -> z. With cheb_m > 1 each block is
    p(B A) B with p the degree-(cheb_m - 1) Chebyshev polynomial of the
    interval [frac b, b], b = 1.05 lambda_max(B A) (Saad, Algorithm 12.1,
    with B in place of the Jacobi diagonal): cheb_m applications of B and
    cheb_m - 1 of A = hdif_gram per block. `lam`: lambda_max per block
    (sigma, u), by default from hx_block_lambdas. `blocks`: prebuilt
    (B_{k-1}, B_k) from hx_blocks. `single`: the Chebyshev steps multiply by
    the FP32 path's symmetric FP32 copy of each block (single_symmetric)."""
    ns, nu = c.ndofs(k - 1), c.ndofs(k)
    hs, hu = blocks or hx_blocks(c, coords, k, n, sweeps)
    if cheb_m <= 1:
        def prec(r):
            parts = [hs.apply(r[:ns])] if ns else []
            parts += [hu.apply(r[ns:ns + nu]), r[ns + nu:]]
            return np.concatenate(parts)
        return prec
    if lam is None:
        lam = hx_block_lambdas(c, coords, k, n, sweeps=sweeps, blocks=(hs, hu), single=single)
    conv = single_symmetric if single else (lambda A: A)
    As = conv(c.hdif_gram(k - 1).to_scipy().tocsr()) if ns else None
    Au = conv(c.hdif_gram(k).to_scipy().tocsr())

    def cheb(B, A, lmax, r):
        b = CHEB_LMAX_MARGIN * lmax
        a = frac * b
        theta, delta = 0.5 * (b + a), 0.5 * (b - a)
        sigma1 = theta / delta
        rho = 1.0 / sigma1
        res = B.apply(r)
        d = res / theta
        z = np.zeros_like(r)
        for _ in range(cheb_m - 1):
            res = res - B.apply(A @ d)
            z = z + d
            rho_next = 1.0 / (2.0 * sigma1 - rho)
            d = rho_next * rho * d + (2.0 * rho_next / delta) * res
            rho = rho_next
        return z + d

    def prec(r):
        parts = [cheb(hs, As, lam[0], r[:ns])] if ns else []
        parts += [cheb(hu, Au, lam[1], r[ns:ns + nu]), r[ns + nu:]]
        return np.concatenate(parts)
    return prec


# ---------------------------------------------------------------------------
# Analytic data: quadrature, source loads, L2 projection and error, de Rham map
# ---------------------------------------------------------------------------
# A field reaches these functions as its values at quadrature nodes, a k-form
# value being its C(dim, k) components on the colex basis dx^I. "reference":
# components in each cell's reference frame (a Section's value,
# derham/src/section.rs:67); "ambient": components in the coordinates of the
# mesh, pulled back cell by cell (Pullback::at, section.rs:182-195).

def _colex_multisets(nsym, degree):
    """Weakly increasing words of length `degree` over range(nsym) in colex
    order (Composition::all, multiindex/src/composition.rs:115-119, through
    Repetition::Allowed.words, multiindex/src/monotone.rs:161-179)."""
    import itertools
    words = list(itertools.combinations_with_replacement(range(nsym), degree))
    words.sort(key=lambda w: tuple(reversed(w)))
    return words


def gm_rule(dim, degree):
    """SimplexQuadRule::degree (simplicial/src/atlas/quadrature.rs:37-83): the
    Grundmann-Moeller rule of index s = degree // 2 on the dim-simplex, exact
    for polynomials of degree 2s + 1; barycentric nodes [np, dim + 1] and
    weights normalized to sum to one."""
    n, s = dim, degree // 2
    d = 2 * s + 1
    pts, ws = [], []
    for i in range(s + 1):
        den = float(d + n - 2 * i)
        sign = 1.0 if MUTATION == 41 else (-1.0) ** i
        w = sign * 2.0 ** (-2 * s) * den ** d / (math.factorial(i) * math.factorial(d + n - i))
        for word in _colex_multisets(n + 1, s - i):
            parts = [0] * (n + 1)
            for sym in word:
                parts[sym] += 1
            pts.append([(2 * b + 1) / den for b in parts])
            ws.append(w)
    ws = np.array(ws)
    total = 0.0
    for w in ws:
        total += w
    return np.array(pts), ws / total


def _subsets(n, k):
    """k-subsets of range(n) in colex order."""
    import itertools
    out = list(itertools.combinations(range(n), k))
    out.sort(key=lambda t: tuple(reversed(t)))
    return out


def _minors(A, rows, cols):
    """[len(rows), len(cols)] determinants of A's minors on the given index
    subsets (Lambda^k of a matrix); A is [..., m, n]."""
    A = np.asarray(A)
    lead = A.shape[:-2]
    out = np.empty(lead + (len(rows), len(cols)))
    for a, R in enumerate(rows):
        for b, Cc in enumerate(cols):
            if len(R) == 0:
                out[..., a, b] = 1.0
            else:
                out[..., a, b] = np.linalg.det(A[..., list(R), :][..., list(Cc)])
    return out


def bary_pullback_table(d, k):
    """P = Lambda^k(Dlambda) [C(d+1, k), C(d, k)]: row J a k-subset of the cell's
    vertices, column I an axis k-subset; Dlambda has row 0 = -1 and row j = e_(j-1)
    (unit_difbarys, simplicial/src/atlas.rs)."""
    Dl = np.zeros((d + 1, d))
    Dl[0, :] = -1.0
    Dl[1:, :] = np.eye(d)
    return _minors(Dl, _subsets(d + 1, k), _subsets(d, k))


def whitney_values(d, k, bary):
    """WhitneyLsf::at_bary (derham/src/interpolate/form.rs:109-114) for every
    DOF of grade k at barycentric points bary [np, d+1]: [np, C(d+1, k+1), C(d, k)],
    W_sigma = k! sum_p (-1)^p lambda_(sigma_p) P[sigma minus sigma_p]."""
    P = bary_pullback_table(d, k)
    vsub = _subsets(d + 1, k)
    rank = {s: i for i, s in enumerate(vsub)}
    dofs = _subsets(d + 1, k + 1)
    out = np.zeros((len(bary), len(dofs), P.shape[1]))
    for i, sig in enumerate(dofs):
        for p, v in enumerate(sig):
            rest = tuple(x for x in sig if x != v)
            kf = 1.0 if MUTATION == 42 else math.factorial(k)
            out[:, i, :] += ((-1.0) ** p * kf) * bary[:, v:v + 1] * P[rank[rest]][None, :]
    return out


def cell_frames(cx: Complex, lengths):
    """Per cell: vol, g^-1 [nc, d, d] (SimplexLengthsSq::metric,
    regge/src/lengths/simplex.rs:308-326; cell_volume, regge/src/lib.rs:26-28)."""
    d = cx.dim
    edges = cx.cell_faces(1)
    s = np.asarray(lengths)[edges]
    loc = local_subsets(d, 1)
    eidx = {(int(a), int(b)): i for i, (a, b) in enumerate(loc)}
    g = np.zeros((len(edges), d, d))
    for i in range(d):
        g[:, i, i] = s[:, eidx[(0, i + 1)]]
        for j in range(i + 1, d):
            v = 0.5 * ((s[:, eidx[(0, i + 1)]] + s[:, eidx[(0, j + 1)]]) - s[:, eidx[(i + 1, j + 1)]])
            g[:, i, j] = g[:, j, i] = v
    det = np.linalg.det(g)
    vol = np.sqrt(np.abs(det)) / math.factorial(d)
    return vol, np.linalg.inv(g)


def multiform_gramian(ginv, k):
    """Lambda^k g^-1 [nc, C(d,k), C(d,k)] (multiform_metric,
    metric/src/tensor.rs:127-131): minors of g^-1 on colex axis subsets."""
    d = ginv.shape[-1]
    sub = _subsets(d, k)
    return _minors(ginv, sub, sub)


def cell_jacobians(cx: Complex, coords):
    """Columns x_(v_(a+1)) - x_(v_0) of each cell's affine map [nc, D, d]
    (the parametrization's linear_transform, section.rs:184-194)."""
    cells = cx.cells()
    X = np.asarray(coords)[cells]
    return np.transpose(X[:, 1:, :] - X[:, :1, :], (0, 2, 1))


def pullback_ambient(J, f, k):
    """Pullback of ambient k-form components f [nc, np, C(D,k)] along J [nc, D, d]:
    f_ref[I] = sum_K det(J[K, I]) f[K] (Tensor::pullback, the contravariant
    action of Lambda^k)."""
    D, d = J.shape[1], J.shape[2]
    if MUTATION == 44:
        J = np.transpose(J, (0, 2, 1))
    PB = _minors(J, _subsets(D, k), _subsets(d, k))  # [nc, C(D,k), C(d,k)]
    return np.einsum("cKI,cnK->cnI", PB, f)


def cell_quadrature_points(cx: Complex, coords, degree):
    """The embedded position of every cell's quadrature nodes [nc, np, D]:
    x = sum_i lambda_i x_(v_i) (chart.point, the MeshPoint the integrand sees)."""
    bary, _ = gm_rule(cx.dim, degree)
    X = np.asarray(coords)[cx.cells()]
    return np.einsum("ni,cia->cna", bary, X)


def source_load(cx: Complex, lengths, k, degree, field, frame="reference", coords=None):
    """SourceForm (formoniq/src/operators.rs:599-635) assembled as a load vector
    (LinearForm assembly, galerkin.rs:205-312): per cell
    b_sigma = vol sum_n w_n <f(x_n), W_sigma(x_n)>_(Lambda^k g^-1), summed into
    the global vector in cell order. field [nc, np, C]."""
    d = cx.dim
    bary, w = gm_rule(d, degree)
    vol, ginv = cell_frames(cx, lengths)
    L = multiform_gramian(ginv, k)
    f = np.asarray(field, dtype=np.float64)
    if frame == "ambient":
        f = pullback_ambient(cell_jacobians(cx, coords), f, k)
    W = whitney_values(d, k, bary)  # [np, nf, NA]
    if MUTATION == 43:
        L = np.broadcast_to(np.eye(L.shape[-1]), L.shape)
    h = np.einsum("cij,cnj->cni", L, f)
    el = np.einsum("n,nsi,cni->cs", w, W, h) * vol[:, None]
    b = np.zeros(cx.nsimplices(k))
    faces = cx.cell_faces(k) if k < d else np.arange(len(vol))[:, None]
    np.add.at(b, faces, el)
    return b


def whitney_at_nodes(cx: Complex, k, cochain, degree):
    """WhitneyInterpolant::eval (derham/src/interpolate/interpolant.rs:55-62) at
    every cell's quadrature nodes, reference frame: [nc, np, C(d, k)]."""
    d = cx.dim
    bary, _ = gm_rule(d, degree)
    W = whitney_values(d, k, bary)
    faces = cx.cell_faces(k) if k < d else np.arange(cx.nsimplices(d))[:, None]
    cv = np.asarray(cochain)[faces]  # [nc, nf]
    return np.einsum("cs,nsi->cni", cv, W)


def fe_l2_error(cx: Complex, lengths, k, cochain, exact, frame="reference", coords=None):
    """fe_l2_error (formoniq/src/fe.rs:50-74): sqrt(sum_K vol_K sum_n w_n
    <e, e>_(Lambda^k g^-1)), e = exact - W c at the degree-3 nodes."""
    d = cx.dim
    _, w = gm_rule(d, 3)
    vol, ginv = cell_frames(cx, lengths)
    L = multiform_gramian(ginv, k)
    f = np.asarray(exact, dtype=np.float64)
    if frame == "ambient":
        f = pullback_ambient(cell_jacobians(cx, coords), f, k)
    e = f - whitney_at_nodes(cx, k, cochain, 3)
    per = np.einsum("n,cni,cij,cnj->c", w, e, L, e) * (1.0 if MUTATION == 46 else vol)
    return math.sqrt(float(np.sum(per)))


def l2_projection(c: WhitneyComplex, k, degree, field, frame="reference", coords=None):
    """l2_projection (formoniq/src/fe.rs:87-123): M_k x = b with b the source
    load, by Jacobi CG to rtol 1e-12 on a Riemannian geometry, else a direct
    solve (the reference's LU)."""
    b = source_load(c.cx, c.lengths, k, degree, field, frame, coords)
    M = c.mass(k)
    _, ginv = cell_frames(c.cx, c.lengths)
    riem = bool(np.all(np.linalg.eigvalsh(ginv) > 0))
    if riem:
        x, rep = cg(M, b, inv_diag=1.0 / M.diagonal(), rtol=1e-12, max_iters=100 * len(b) + 100)
        return x
    import scipy.sparse.linalg as spla
    return spla.spsolve(M.to_scipy().tocsc(), b)


def simplex_quadrature_points(cx: Complex, coords, k, degree):
    """Positions of the quadrature nodes of every k-simplex [N_k, np, D]."""
    bary, _ = gm_rule(k, degree)
    X = np.asarray(coords)[cx.skeleton(k)]
    return np.einsum("ni,sia->sna", bary, X)


def derham_map(cx: Complex, coords, k, degree, field):
    """derham_map (derham/src/project.rs:55-96): c_s = (1/k!) sum_n w_n
    <f(x_n), e_1 ^ ... ^ e_k>, the pairing of the ambient k-form with the
    tangent blade of the simplex (edge vectors e_a = x_(s_a) - x_(s_0)), by the
    degree-`degree` rule on the k-simplex. field [N_k, np, C(D, k)]."""
    D = np.asarray(coords).shape[1]
    _, w = gm_rule(k, degree)
    X = np.asarray(coords)[cx.skeleton(k)]
    E = np.transpose(X[:, 1:, :] - X[:, :1, :], (0, 2, 1))  # [N_k, D, k]
    blade = _minors(E, _subsets(D, k), [tuple(range(k))])[..., 0]  # [N_k, C(D,k)]
    integrand = np.einsum("snK,sK->sn", np.asarray(field, dtype=np.float64), blade)
    return (integrand @ w) / (1 if MUTATION == 45 else math.factorial(k))
