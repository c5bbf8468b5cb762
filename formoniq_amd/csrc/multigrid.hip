// Mesh-independent block preconditioner of the mixed KKT system: geometric
// multigrid on Kuhn grids inside the Hiptmair-Xu auxiliary-space
// preconditioner, applied on the device within the MINRES iterations.
//
// Reference behaviour (luiswirth/formoniq, crates/):
//   VCycle::cycle                    iterative/src/multigrid.rs (x = 0, `sweeps`
//                                    smoothing sweeps, residual, restrict, coarse
//                                    correction, prolong, `sweeps` sweeps; direct
//                                    solve on the coarsest level)
//   Jacobi::smoother                 iterative/src/precond.rs:89-94 (omega = 2/3)
//   stationary::sweeps               iterative/src/stationary.rs:61-74
//   AuxiliarySpace                   iterative/src/aux_space.rs (B = S + sum Pi B_i Pi^T)
//   hx_preconditioner                formoniq/src/hx.rs:244-275
//   vector_nodal_prolongation        formoniq/src/hx.rs:86-112
//   RefinementTower::grade_vcycle    formoniq/src/multigrid.rs:115-138
//   mixed_block_preconditioner       formoniq/src/problems/elliptic.rs:29-47
//
// The block preconditioner of the KKT system of grade k is diag(B_{k-1}, B_k, I)
// with B_j the HX preconditioner of hdif_gram(j) = M_j + D_j^T M_{j+1} D_j:
//   B_j = S_j + D_{j-1} B_{j-1} D_{j-1}^T + Pi_j (V (x) I) Pi_j^T,    B_0 = S_0 + V,
// S_j = 2/3 diag(hdif_gram(j))^-1, Pi_j the vector-nodal interpolation of grade
// j (Pi_0 = I) and V one V-cycle of the nodal problem hdif_gram(0), the same
// cycle for every copy. Since D_{j-1} D_{j-2} = 0, the recursion ends one grade
// down: D_{j-1} B_{j-1} D_{j-1}^T = D_{j-1} (S_{j-1} + Pi_{j-1} V Pi_{j-1}^T) D_{j-1}^T.
// Every V of one application is the same operator, so all of them run as one
// multi-right-hand-side cycle ("slots", slot-major vectors of the nodal size).
//
// Hierarchy. The reference refines a coarse mesh into a tower. Here the fine
// mesh is given, and when it is the Kuhn triangulation of the n^d box (vertex
// c at sum_a c_a (n+1)^a, as feec_kuhn_grid numbers it) the tower is recovered
// by coarsening: the grid of m = ceil(n/2) cells per axis, on which fine vertex
// 2w + delta (delta in {0,1}^d) is the midpoint of the coarse Kuhn edge
// (w, w + delta), so the nodal prolongation takes 1 or 1/2, 1/2 of coarse
// values (for odd n the coarse grid overhangs the box by one fine cell; every
// coarse vertex still has a fine child). The coarse operators are the Galerkin
// products A_c = P^T A P (the reference's multigrid.rs module note), so no
// coarse geometry is needed and jittered fine meshes are covered. Levels are
// added while a level has more than kCoarseMax unknowns; the coarsest is
// solved by its dense inverse, formed on the host by Cholesky.
//
// Precision. The FP64 path runs every piece in FP64; the mixed paths in FP32
// (the FP64_PC32 solve rounds r to FP32 on entry and widens z on exit).
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "enqueue.hpp"
#include "feec_internal.hpp"

namespace feec {

namespace {

constexpr int TPB = 256;
constexpr int NBLK = 4096;
constexpr int MAXNV = 16;           // slots of one multi-right-hand-side cycle
constexpr int64_t kCoarseMax = 200;  // largest level solved directly
constexpr double kOmega = 2.0 / 3.0;

inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + TPB - 1) / TPB); }
inline unsigned red_grid(int64_t n) {
  return (unsigned)std::min<int64_t>(NBLK, std::max<int64_t>(1, (n + TPB - 1) / TPB));
}

__device__ __forceinline__ bool skip(const MinresS* S) { return S && S->done; }

// ---------------------------------------------------------------------------
// multi-vector kernels: nv slots of length n, slot-major (v * n + i)
// ---------------------------------------------------------------------------
template <class TP>
__global__ void __launch_bounds__(TPB) k_mg_scale(const TP* __restrict__ dinv, const TP* __restrict__ r,
                                                  TP* __restrict__ x, int64_t n, int nv,
                                                  const MinresS* S) {
  if (skip(S)) return;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const TP d = dinv[i];
  for (int v = 0; v < nv; ++v) x[v * n + i] = d * r[v * n + i];
}

// xo = x + dinv (r - A x)   (mode 0: one Jacobi sweep)
// e  = r - A x               (mode 1: residual)
template <class TP>
__global__ void __launch_bounds__(TPB) k_mg_rowop(const int64_t* __restrict__ rp,
                                                  const int32_t* __restrict__ col,
                                                  const TP* __restrict__ val,
                                                  const TP* __restrict__ dinv,
                                                  const TP* __restrict__ r, const TP* __restrict__ x,
                                                  TP* __restrict__ out, int64_t n, int nv, int mode,
                                                  const MinresS* S) {
  if (skip(S)) return;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  TP acc[MAXNV];
#pragma unroll
  for (int v = 0; v < MAXNV; ++v) acc[v] = TP(0);
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
    const int64_t c = col[p];
    const TP a = val[p];
#pragma unroll
    for (int v = 0; v < MAXNV; ++v)
      if (v < nv) acc[v] += a * x[v * n + c];
  }
  const TP d = mode == 0 ? dinv[i] : TP(0);
#pragma unroll
  for (int v = 0; v < MAXNV; ++v) {
    if (v >= nv) break;
    const TP res = r[v * n + i] - acc[v];
    out[v * n + i] = mode == 0 ? x[v * n + i] + d * res : res;
  }
}

// y = M x (add = 0) or y += M x (add = 1), M an (nr x nc) CSR, x of stride nc.
template <class TP>
__global__ void __launch_bounds__(TPB) k_mg_spmv(const int64_t* __restrict__ rp,
                                                 const int32_t* __restrict__ col,
                                                 const TP* __restrict__ val, const TP* __restrict__ x,
                                                 TP* __restrict__ y, int64_t nr, int64_t nc, int nv,
                                                 int add, const MinresS* S) {
  if (skip(S)) return;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nr) return;
  TP acc[MAXNV];
#pragma unroll
  for (int v = 0; v < MAXNV; ++v) acc[v] = TP(0);
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
    const int64_t c = col[p];
    const TP a = val[p];
#pragma unroll
    for (int v = 0; v < MAXNV; ++v)
      if (v < nv) acc[v] += a * x[v * nc + c];
  }
#pragma unroll
  for (int v = 0; v < MAXNV; ++v) {
    if (v >= nv) break;
    y[v * nr + i] = add ? y[v * nr + i] + acc[v] : acc[v];
  }
}

// x = C r, C dense (n x n, row-major): one block per row, all slots.
template <class TP>
__global__ void __launch_bounds__(TPB) k_mg_dense(const TP* __restrict__ C, const TP* __restrict__ r,
                                                  TP* __restrict__ x, int64_t n, int nv,
                                                  const MinresS* S) {
  if (skip(S)) return;
  __shared__ TP sh[TPB / 64][MAXNV];
  const int64_t i = blockIdx.x;
  TP acc[MAXNV];
#pragma unroll
  for (int v = 0; v < MAXNV; ++v) acc[v] = TP(0);
  for (int64_t j = threadIdx.x; j < n; j += blockDim.x) {
    const TP c = C[i * n + j];
#pragma unroll
    for (int v = 0; v < MAXNV; ++v)
      if (v < nv) acc[v] += c * r[v * n + j];
  }
#pragma unroll
  for (int v = 0; v < MAXNV; ++v) {
    TP t = acc[v];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    acc[v] = t;
  }
  if ((threadIdx.x & 63) == 0)
    for (int v = 0; v < MAXNV; ++v) sh[threadIdx.x >> 6][v] = acc[v];
  __syncthreads();
  if (threadIdx.x < (unsigned)nv) {
    TP t = TP(0);
    for (int w = 0; w < TPB / 64; ++w) t += sh[w][threadIdx.x];
    x[threadIdx.x * n + i] = t;
  }
}

// ---------------------------------------------------------------------------
// HX pieces on the KKT vector
// ---------------------------------------------------------------------------
// z = s r on the preconditioned rows (s = 2/3 inv diag), z = r on the border rows.
template <class TIN, class TP>
__global__ void k_hx_init(const TIN* __restrict__ r, const TP* __restrict__ s, TP* __restrict__ z,
                          int64_t n, int64_t N, const MinresS* S) {
  if (skip(S)) return;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= N) return;
  z[i] = i < n ? s[i] * (TP)r[i] : (TP)r[i];
}

// y[v * nr + row] = sum_p val[p * C + v] x[col[p]] for v < C: the C components
// of a transposed vector-nodal interpolation (or, with C = 1, any CSR).
template <class TX, class TP>
__global__ void k_hx_gather(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                            const TP* __restrict__ val, int C, const TX* __restrict__ x,
                            TP* __restrict__ y, int64_t nr, const MinresS* S) {
  if (skip(S)) return;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nr) return;
  TP acc[MAXNV];
#pragma unroll
  for (int v = 0; v < MAXNV; ++v) acc[v] = TP(0);
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
    const TP xv = (TP)x[col[p]];
#pragma unroll
    for (int v = 0; v < MAXNV; ++v)
      if (v < C) acc[v] += val[p * C + v] * xv;
  }
  for (int v = 0; v < C; ++v) y[v * nr + i] = acc[v];
}

// y[row] = (a ? a[row] b[row] : y[row] if add else 0) + sum_p sum_v val[p * C + v] x[v * nx + col[p]]
template <class TP>
__global__ void k_hx_scatter(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                             const TP* __restrict__ val, int C, const TP* __restrict__ x, int64_t nx,
                             const TP* __restrict__ a, const TP* __restrict__ b,
                             TP* __restrict__ y, int64_t nr, int add, const MinresS* S) {
  if (skip(S)) return;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nr) return;
  TP acc = TP(0);
  if (rp)
    for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
      const int64_t c = col[p];
      for (int v = 0; v < C; ++v) acc += val[p * C + v] * x[v * nx + c];
    }
  if (a)
    acc += a[i] * b[i];
  else if (add)
    acc += y[i];
  y[i] = acc;
}

// slot = r (TIN -> TP), or y += x (elementwise) for the grade-0 V terms.
template <class TIN, class TP>
__global__ void k_hx_copy(const TIN* __restrict__ x, TP* __restrict__ y, int64_t n, int add,
                          const MinresS* S) {
  if (skip(S)) return;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  y[i] = add ? y[i] + (TP)x[i] : (TP)x[i];
}

__device__ __forceinline__ double dev_block_sum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (unsigned w = 0; w < blockDim.x / 64; ++w) r += sh[w];
  return r;
}

// z (TOUT) = zt (TP) and the partial sums of r . z (red_grid(N) blocks).
template <class TIN, class TP, class TOUT>
__global__ void __launch_bounds__(TPB) k_hx_final(const TIN* __restrict__ r, const TP* __restrict__ zt,
                                                  TOUT* __restrict__ z, int64_t N, const MinresS* S,
                                                  double* __restrict__ part) {
  if (skip(S)) return;
  __shared__ double sh[TPB / 64];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    const TOUT v = (TOUT)zt[i];
    z[i] = v;
    acc += (double)r[i] * (double)v;
  }
  const double t = dev_block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <class TS, class TD>
__global__ void k_convert(const TS* __restrict__ a, TD* __restrict__ b, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (TD)a[i];
}

__global__ void k_scaled_inverse_diag(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                                      const double* __restrict__ val, int64_t n, double omega,
                                      double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p)
    if (col[p] == i) d = val[p];
  out[i] = d != 0.0 ? omega / d : 0.0;
}

__global__ void k_scale(const double* __restrict__ a, int64_t n, double s, double* __restrict__ b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = s * a[i];
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
struct HostCsr {
  int64_t nrows = 0, ncols = 0;
  std::vector<int64_t> rp;
  std::vector<int32_t> col;
  std::vector<double> val;  // nnz * C
  int C = 1;
};

void upload(Ctx* ctx, const HostCsr& h, Csr& out) {
  out.ctx = ctx;
  out.nrows = h.nrows;
  out.ncols = h.ncols;
  out.nnz = (int64_t)h.col.size();
  out.row_ptr.alloc(ctx, h.nrows + 1);
  out.col.alloc(ctx, std::max<int64_t>(out.nnz, 1));
  out.val.alloc(ctx, std::max<size_t>(h.val.size(), 1));
  copy_to_device(ctx, out.row_ptr.p, h.rp.data(), (h.nrows + 1) * 8);
  if (out.nnz) copy_to_device(ctx, out.col.p, h.col.data(), out.nnz * 4);
  if (!h.val.empty()) copy_to_device(ctx, out.val.p, h.val.data(), h.val.size() * 8);
}

// Transpose of a host CSR with C values per entry.
HostCsr transpose(const HostCsr& a) {
  HostCsr t;
  t.nrows = a.ncols;
  t.ncols = a.nrows;
  t.C = a.C;
  t.rp.assign(t.nrows + 1, 0);
  for (int32_t c : a.col) ++t.rp[c + 1];
  for (int64_t i = 0; i < t.nrows; ++i) t.rp[i + 1] += t.rp[i];
  t.col.resize(a.col.size());
  t.val.resize(a.val.size());
  std::vector<int64_t> at(t.rp.begin(), t.rp.end() - 1);
  for (int64_t r = 0; r < a.nrows; ++r)
    for (int64_t p = a.rp[r]; p < a.rp[r + 1]; ++p) {
      const int64_t q = at[a.col[p]]++;
      t.col[q] = (int32_t)r;
      for (int v = 0; v < a.C; ++v) t.val[q * a.C + v] = a.val[p * a.C + v];
    }
  return t;
}

int64_t ipow(int64_t b, int e) {
  int64_t r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}

// Nodal prolongation of the Kuhn grid of g cells per axis from the grid of
// ceil(g/2): fine vertex 2w + delta takes 1 (delta = 0) or 1/2, 1/2 of the
// coarse vertices w and w + delta. With `interior`, rows and columns are the
// interior vertices (all coordinates in 1..g-1 / coarse 2w <= g - 1), numbered
// in vertex order, and entries on other coarse vertices are dropped.
HostCsr kuhn_prolongation(int d, int g, bool interior) {
  const int gc = (g + 1) / 2;
  const int64_t nf = ipow(g + 1, d), nc = ipow(gc + 1, d);
  std::vector<int32_t> fmap(nf, -1), cmap(nc, -1);
  int64_t nfi = 0, nci = 0;
  for (int64_t v = 0; v < nf; ++v) {
    bool in = true;
    int64_t r = v;
    for (int a = 0; a < d; ++a, r /= (g + 1)) {
      const int64_t c = r % (g + 1);
      in = in && c > 0 && c < g;
    }
    if (!interior || in) fmap[v] = (int32_t)nfi++;
  }
  for (int64_t w = 0; w < nc; ++w) {
    bool in = true;
    int64_t r = w;
    for (int a = 0; a < d; ++a, r /= (gc + 1)) {
      const int64_t c = r % (gc + 1);
      in = in && c > 0 && 2 * c <= g - 1;
    }
    if (!interior || in) cmap[w] = (int32_t)nci++;
  }
  HostCsr P;
  P.nrows = nfi;
  P.ncols = nci;
  P.rp.assign(nfi + 1, 0);
  P.col.reserve(nfi * 2);
  P.val.reserve(nfi * 2);
  int64_t row = 0;
  for (int64_t v = 0; v < nf; ++v) {
    if (fmap[v] < 0) continue;
    int64_t lo = 0, hi = 0, stride = 1, r = v;
    bool odd = false;
    for (int a = 0; a < d; ++a, r /= (g + 1), stride *= (gc + 1)) {
      const int64_t c = r % (g + 1), dl = c & 1;
      lo += (c >> 1) * stride;
      hi += ((c >> 1) + dl) * stride;
      odd = odd || dl;
    }
    auto push = [&](int64_t w, double x) {
      if (cmap[w] >= 0) {
        P.col.push_back(cmap[w]);
        P.val.push_back(x);
      }
    };
    if (!odd) {
      push(lo, 1.0);
    } else if (lo < hi) {
      push(lo, 0.5);
      push(hi, 0.5);
    }
    P.rp[++row] = (int64_t)P.col.size();
  }
  return P;
}

// Dense inverse of an SPD matrix by Cholesky; false if it is not SPD.
bool spd_inverse(int64_t n, std::vector<double>& a) {
  std::vector<double> L(n * n, 0.0);
  for (int64_t j = 0; j < n; ++j) {
    double s = a[j * n + j];
    for (int64_t k = 0; k < j; ++k) s -= L[j * n + k] * L[j * n + k];
    if (!(s > 0.0)) return false;
    const double ljj = std::sqrt(s);
    L[j * n + j] = ljj;
    for (int64_t i = j + 1; i < n; ++i) {
      double t = a[i * n + j];
      for (int64_t k = 0; k < j; ++k) t -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = t / ljj;
    }
  }
  // inverse column by column: L L^T x = e_c
  std::vector<double> inv(n * n), y(n);
  for (int64_t c = 0; c < n; ++c) {
    for (int64_t i = 0; i < n; ++i) {
      double t = i == c ? 1.0 : 0.0;
      for (int64_t k = 0; k < i; ++k) t -= L[i * n + k] * y[k];
      y[i] = t / L[i * n + i];
    }
    for (int64_t i = n - 1; i >= 0; --i) {
      double t = y[i];
      for (int64_t k = i + 1; k < n; ++k) t -= L[k * n + i] * inv[k * n + c];
      inv[i * n + c] = t / L[i * n + i];
    }
  }
  a.swap(inv);
  return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// Kuhn grid detection
// ---------------------------------------------------------------------------
int kuhn_grid_n(Mesh& m) {
  std::lock_guard<std::mutex> lock(m.mu);
  if (m.kuhn_n >= 0) return m.kuhn_n;
  m.kuhn_n = 0;
  const int d = m.d;
  const int64_t nv = m.n[0], nc = m.n[d];
  int n = (int)std::llround(std::pow((double)nv, 1.0 / d)) - 1;
  int64_t fact = 1;
  for (int i = 2; i <= d; ++i) fact *= i;
  if (n < 1 || ipow(n + 1, d) != nv || fact * ipow(n, d) != nc) return 0;
  std::vector<int32_t> cv((size_t)nc * (d + 1));
  FEEC_HIP(hipMemcpyAsync(cv.data(), m.faces[0].p, cv.size() * 4, hipMemcpyDeviceToHost, m.ctx->stream));
  m.ctx->sync();
  int64_t stride[5];
  stride[0] = 1;
  for (int a = 1; a <= d; ++a) stride[a] = stride[a - 1] * (n + 1);
  for (int64_t c = 0; c < nc; ++c) {
    const int32_t* v = &cv[c * (d + 1)];
    int64_t r = v[0];
    for (int a = 0; a < d; ++a, r /= (n + 1))
      if (r % (n + 1) >= n) return 0;  // the origin of a box
    unsigned used = 0;
    for (int i = 0; i < d; ++i) {
      const int64_t diff = (int64_t)v[i + 1] - v[i];
      int axis = -1;
      for (int a = 0; a < d; ++a)
        if (diff == stride[a]) axis = a;
      if (axis < 0 || (used & (1u << axis))) return 0;
      used |= 1u << axis;
    }
  }
  // distinct cells (the mesh deduplicates), each a Kuhn simplex, d! n^d of them: all of them
  m.kuhn_n = n;
  return n;
}

// ---------------------------------------------------------------------------
// the preconditioner object
// ---------------------------------------------------------------------------
struct MgLevel {
  int grid = 0;  // cells per axis
  int64_t n = 0;
  Csr A;                  // FP64 operator
  DBuf<float> A32;
  DBuf<double> dinv;      // omega / diag(A)
  DBuf<float> dinv32;
  Csr P, R;               // from the next coarser level (n x n_c) and its transpose
  DBuf<float> P32, R32;
  // slot vectors: rhs, iterate (two), residual
  DBuf<double> r64, x64, y64, e64;
  DBuf<float> r32, x32, y32, e32;
};

struct HxBlock {
  int j = -1;           // grade of the block
  int64_t off = 0, n = 0;
  // gradient part (j >= 1): D_{j-1} and D_{j-1}^T, and the smoother of grade j-1
  Csr D, DT;
  DBuf<float> D32, DT32;
  DBuf<double> sl;      // 2/3 inv diag hdif_gram(j-1), on the (j-1)-DOFs
  DBuf<float> sl32;
  DBuf<double> t64, u64;  // D^T r and the grade-(j-1) correction
  DBuf<float> t32, u32;
  // vector-nodal interpolations: of grade j (j >= 1) and of grade j-1 (j-1 >= 1)
  int C = 0, Cl = 0;
  Csr Pi, PiT, Pil, PilT;  // values C (Cl) per entry
  DBuf<float> Pi32, PiT32, Pil32, PilT32;
  int slot_v = -1;      // first slot of grade j's V terms (or the single slot at j = 0)
  int slot_t = -1;      // first slot of the gradient's V terms
};

struct Hx {
  Ctx* ctx = nullptr;
  int d = 0, k = 0, bc = 0, grid = 0, sweeps = 2;
  int64_t n = 0, N = 0, nv_nodes = 0;
  int nv = 0;  // slots
  std::vector<MgLevel> L;  // finest first; the last is solved densely
  DBuf<double> cinv;
  DBuf<float> cinv32;
  std::vector<HxBlock> blocks;
  DBuf<double> s64, zt64;  // KKT-vector smoother diagonal and work vector
  DBuf<float> s32, zt32;
  int64_t bytes32 = 0, launches = 0;
};

void hx_free(Hx* h) { delete h; }
int64_t hx_bytes_per_apply(const Hx& h) { return h.bytes32; }
int hx_launches_per_apply(const Hx& h) { return (int)h.launches; }

namespace {

// Vector-nodal interpolation of grade j (formoniq/src/hx.rs:86-112): row of
// the j-simplex sigma, entry per vertex a of sigma with C = C(d, j) values,
// value I = (1/(j+1)!) det(E_I), E the edge vectors of sigma from its first
// vertex and E_I their components I (colex j-subsets of the d axes).
// Rows and columns restricted to the complex's DOFs (relative: interior).
HostCsr vector_nodal(Mesh& m, const std::vector<double>& coords, int j, const int32_t* rowmap,
                     const int32_t* colmap, int64_t nrows, int64_t ncols) {
  const int d = m.d;
  const int64_t ns = m.n[j];
  std::vector<int64_t> sv(ns * (j + 1));
  if (feec_mesh_skeleton(static_cast<feec_mesh*>(&m), j, sv.data()) != FEEC_OK)
    fail(FEEC_E_DEVICE, feec_last_error());
  std::vector<std::vector<int>> subsets;
  for (unsigned mask = 0; mask < (1u << d); ++mask)
    if (__builtin_popcount(mask) == j) {
      std::vector<int> s;
      for (int a = 0; a < d; ++a)
        if (mask & (1u << a)) s.push_back(a);
      subsets.push_back(s);
    }
  // colex order of subsets = increasing mask
  const int C = (int)subsets.size();
  double fact = 1.0;
  for (int i = 2; i <= j + 1; ++i) fact *= i;
  const double factor = 1.0 / fact;
  HostCsr Pi;
  Pi.nrows = nrows;
  Pi.ncols = ncols;
  Pi.C = C;
  Pi.rp.assign(nrows + 1, 0);
  std::vector<int64_t> order(ns, -1);
  for (int64_t s = 0; s < ns; ++s) {
    const int32_t r = rowmap ? rowmap[s] : (int32_t)s;
    if (r >= 0) order[r] = s;
  }
  for (int64_t r = 0; r < nrows; ++r) {
    const int64_t s = order[r];
    const int64_t* v = &sv[s * (j + 1)];
    double E[4][4];
    for (int i = 0; i < j; ++i)
      for (int a = 0; a < d; ++a) E[i][a] = coords[v[i + 1] * d + a] - coords[v[0] * d + a];
    double w[6];
    for (int I = 0; I < C; ++I) {
      const auto& ax = subsets[I];
      double det;
      if (j == 1) {
        det = E[0][ax[0]];
      } else if (j == 2) {
        det = E[0][ax[0]] * E[1][ax[1]] - E[0][ax[1]] * E[1][ax[0]];
      } else {  // j == 3
        det = E[0][ax[0]] * (E[1][ax[1]] * E[2][ax[2]] - E[1][ax[2]] * E[2][ax[1]]) -
              E[0][ax[1]] * (E[1][ax[0]] * E[2][ax[2]] - E[1][ax[2]] * E[2][ax[0]]) +
              E[0][ax[2]] * (E[1][ax[0]] * E[2][ax[1]] - E[1][ax[1]] * E[2][ax[0]]);
      }
      w[I] = factor * det;
    }
    // entries sorted by column
    std::vector<std::pair<int32_t, int>> ents;
    for (int p = 0; p <= j; ++p) {
      const int32_t c = colmap ? colmap[v[p]] : (int32_t)v[p];
      if (c >= 0) ents.push_back({c, p});
    }
    std::sort(ents.begin(), ents.end());
    for (auto& e : ents) {
      Pi.col.push_back(e.first);
      for (int I = 0; I < C; ++I) Pi.val.push_back(w[I]);
    }
    Pi.rp[r + 1] = (int64_t)Pi.col.size();
  }
  return Pi;
}

HostCsr host_of(Ctx* ctx, const Csr& a) {
  HostCsr h;
  h.nrows = a.nrows;
  h.ncols = a.ncols;
  h.rp.resize(a.nrows + 1);
  h.col.resize(a.nnz);
  h.val.resize(a.nnz);
  copy_to_host(ctx, h.rp.data(), a.row_ptr.p, (a.nrows + 1) * 8);
  if (a.nnz) {
    copy_to_host(ctx, h.col.data(), a.col.p, a.nnz * 4);
    copy_to_host(ctx, h.val.data(), a.val.p, a.nnz * 8);
  }
  return h;
}

void to32(Ctx* ctx, const double* a, int64_t n, DBuf<float>& out) {
  if ((int64_t)out.n != std::max<int64_t>(n, 1)) out.alloc(ctx, std::max<int64_t>(n, 1));
  if (n) k_convert<double, float><<<grid_for(n), TPB, 0, ctx->stream>>>(a, out.p, n);
  FEEC_HIP(hipGetLastError());
}

// The nodal operator's block of the KKT preconditioner matrix P (rows and
// columns lo..hi-1), or hdif_gram(0) assembled on its own.
void nodal_operator(Geom& g, int bc, const KktSystem& s, int k, Csr& out) {
  Mesh& m = *g.mesh;
  if (k <= 1 && s.P.nnz) {
    const int64_t lo = k == 1 ? 0 : s.split, hi = k == 1 ? s.split : s.n;
    std::vector<int32_t> map(s.n, -1);
    for (int64_t i = lo; i < hi; ++i) map[i] = (int32_t)(i - lo);
    DBuf<int32_t> dm(m.ctx, std::max<int64_t>(s.n, 1));
    copy_to_device(m.ctx, dm.p, map.data(), s.n * 4);
    csr_select(s.P, dm.p, hi - lo, dm.p, hi - lo, out);
    return;
  }
  assemble_form(g, bc, FEEC_HDIF_GRAM, 0, out);
}

void level_refresh_derived(Ctx* ctx, MgLevel& l) {
  l.dinv.alloc(ctx, std::max<int64_t>(l.n, 1));
  if (l.n)
    k_scaled_inverse_diag<<<grid_for(l.n), TPB, 0, ctx->stream>>>(l.A.row_ptr.p, l.A.col.p, l.A.val.p,
                                                                   l.n, kOmega, l.dinv.p);
  FEEC_HIP(hipGetLastError());
  to32(ctx, l.dinv.p, l.n, l.dinv32);
  to32(ctx, l.A.val.p, l.A.nnz, l.A32);
}

// Galerkin operators and the coarse inverse from the finest operator's values.
// Pattern buffers stay where they are, so captured graphs stay valid.
void hierarchy_values(Hx& h) {
  Ctx* ctx = h.ctx;
  for (size_t i = 0; i + 1 < h.L.size(); ++i) {
    MgLevel& f = h.L[i];
    MgLevel& c = h.L[i + 1];
    Csr AP, RAP;
    csr_mul(f.A, f.P, AP);
    csr_mul(f.R, AP, RAP);
    if (!c.A.row_ptr.p) {
      c.A = std::move(RAP);
    } else {
      FEEC_REQUIRE(RAP.nnz == c.A.nnz, FEEC_E_DEVICE, "coarse operator pattern changed");
      FEEC_HIP(hipMemcpyAsync(c.A.val.p, RAP.val.p, RAP.nnz * 8, hipMemcpyDeviceToDevice, ctx->stream));
      ctx->sync();
    }
    level_refresh_derived(ctx, c);
  }
  MgLevel& last = h.L.back();
  HostCsr a = host_of(ctx, last.A);
  const int64_t n = last.n;
  std::vector<double> dense(n * n, 0.0);
  for (int64_t r = 0; r < n; ++r)
    for (int64_t p = a.rp[r]; p < a.rp[r + 1]; ++p) dense[r * n + a.col[p]] += a.val[p];
  FEEC_REQUIRE(spd_inverse(n, dense), FEEC_E_NOT_SPD, "coarsest multigrid operator is not SPD");
  h.cinv.alloc(ctx, std::max<int64_t>(n * n, 1));
  copy_to_device(ctx, h.cinv.p, dense.data(), n * n * 8);
  to32(ctx, h.cinv.p, n * n, h.cinv32);
  ctx->sync();
}

void alloc_slots(Ctx* ctx, Hx& h) {
  for (auto& l : h.L) {
    const int64_t m = std::max<int64_t>(l.n * h.nv, 1);
    l.r64.alloc(ctx, m);
    l.x64.alloc(ctx, m);
    l.y64.alloc(ctx, m);
    l.e64.alloc(ctx, m);
    l.r32.alloc(ctx, m);
    l.x32.alloc(ctx, m);
    l.y32.alloc(ctx, m);
    l.e32.alloc(ctx, m);
  }
}

void smoother_diag(Hx& h, const KktSystem& s) {
  Ctx* ctx = h.ctx;
  h.s64.alloc(ctx, std::max<int64_t>(h.n, 1));
  if (h.n) k_scale<<<grid_for(h.n), TPB, 0, ctx->stream>>>(s.inv_diag.p, h.n, kOmega, h.s64.p);
  FEEC_HIP(hipGetLastError());
  to32(ctx, h.s64.p, h.n, h.s32);
  for (auto& b : h.blocks) {
    if (b.j < 1) continue;
    // grade j-1 smoother: the sigma block's when it is the other block, else the nodal one
    const HxBlock* lower = nullptr;
    for (auto& o : h.blocks)
      if (o.j == b.j - 1) lower = &o;
    const int64_t nl = b.D.ncols;
    b.sl.alloc(ctx, std::max<int64_t>(nl, 1));
    if (lower)
      FEEC_HIP(hipMemcpyAsync(b.sl.p, h.s64.p + lower->off, nl * 8, hipMemcpyDeviceToDevice, ctx->stream));
    else if (b.j - 1 == 0)
      FEEC_HIP(hipMemcpyAsync(b.sl.p, h.L[0].dinv.p, nl * 8, hipMemcpyDeviceToDevice, ctx->stream));
    to32(ctx, b.sl.p, nl, b.sl32);
  }
}

}  // namespace

Hx* hx_create(Geom& g, int bc, int k, const KktSystem& s, int sweeps, std::string* why) {
  Mesh& m = *g.mesh;
  Ctx* ctx = m.ctx;
  auto no = [&](const std::string& msg) -> Hx* {
    if (why) *why = msg;
    return nullptr;
  };
  if (g.coords.empty()) return no("the geometry has no vertex coordinates (vector-nodal space)");
  if (!g.euclidean) return no("the ambient metric is not Euclidean (HX needs an SPD problem)");
  if (k > 2) return no("grades above 2 are not covered");
  if (s.nh && !(bc == FEEC_FULL && k == 0) && !(bc == FEEC_RELATIVE && k == m.d))
    return no("unexpected harmonic border");
  const int n = kuhn_grid_n(m);
  if (n < 2) return no("the mesh is not a Kuhn grid of the library's numbering");
  auto h = std::make_unique<Hx>();
  h->ctx = ctx;
  h->d = m.d;
  h->k = k;
  h->bc = bc;
  h->grid = n;
  h->sweeps = std::max(1, sweeps);
  h->n = s.n;
  h->N = s.n + s.nh;
  const bool rel = bc == FEEC_RELATIVE;
  if (rel) m.ensure_interior();
  // the nodal hierarchy
  {
    MgLevel l0;
    l0.grid = n;
    nodal_operator(g, bc, s, k, l0.A);
    l0.n = l0.A.nrows;
    h->nv_nodes = l0.n;
    level_refresh_derived(ctx, l0);
    h->L.push_back(std::move(l0));
    int gr = n;
    while (h->L.back().n > kCoarseMax && gr >= 2) {
      MgLevel& f = h->L.back();
      HostCsr P = kuhn_prolongation(m.d, gr, rel);
      FEEC_REQUIRE(P.nrows == f.n, FEEC_E_DEVICE, "prolongation does not match the level");
      HostCsr R = transpose(P);
      upload(ctx, P, f.P);
      upload(ctx, R, f.R);
      to32(ctx, f.P.val.p, f.P.nnz, f.P32);
      to32(ctx, f.R.val.p, f.R.nnz, f.R32);
      MgLevel c;
      gr = (gr + 1) / 2;
      c.grid = gr;
      c.n = P.ncols;
      h->L.push_back(std::move(c));
    }
    if (h->L.back().n > 4 * kCoarseMax) return no("the coarsest multigrid level is too large");
  }
  try {
    hierarchy_values(*h);
  } catch (const Error& e) {
    return no(e.what());
  }
  // blocks: sigma (grade k-1, when k >= 1) and u (grade k)
  const int64_t ns = s.split;
  std::vector<int> grades;
  if (k >= 1) grades.push_back(k - 1);
  grades.push_back(k);
  int slot = 0;
  for (int j : grades) {
    HxBlock b;
    b.j = j;
    b.off = (j == k) ? ns : 0;
    b.n = (j == k) ? s.n - ns : ns;
    const int32_t* map_j = nullptr;
    std::vector<int32_t> hmap_j, hmap_l, hmap_0;
    auto host_map = [&](int grade, std::vector<int32_t>& out) -> const int32_t* {
      if (!rel) return nullptr;
      out.resize(m.n[grade]);
      copy_to_host(ctx, out.data(), m.interior[grade].p, m.n[grade] * 4);
      return out.data();
    };
    map_j = host_map(j, hmap_j);
    const int32_t* map_0 = host_map(0, hmap_0);
    if (j == 0) {
      b.slot_v = slot++;
    } else {
      // D_{j-1} on the complex's DOFs and its transpose
      Csr Dfull;
      mesh_coboundary(m, j - 1, Dfull);
      if (rel)
        csr_select(Dfull, m.interior[j].p, m.n_interior[j], m.interior[j - 1].p, m.n_interior[j - 1],
                   b.D);
      else
        b.D = std::move(Dfull);
      csr_transpose(b.D, b.DT);
      to32(ctx, b.D.val.p, b.D.nnz, b.D32);
      to32(ctx, b.DT.val.p, b.DT.nnz, b.DT32);
      const int64_t nl = b.D.ncols;
      b.t64.alloc(ctx, std::max<int64_t>(nl, 1));
      b.u64.alloc(ctx, std::max<int64_t>(nl, 1));
      b.t32.alloc(ctx, std::max<int64_t>(nl, 1));
      b.u32.alloc(ctx, std::max<int64_t>(nl, 1));
      // vector-nodal of grade j
      HostCsr Pi = vector_nodal(m, g.coords, j, map_j, map_0, b.n, h->nv_nodes);
      HostCsr PiT = transpose(Pi);
      b.C = Pi.C;
      upload(ctx, Pi, b.Pi);
      upload(ctx, PiT, b.PiT);
      to32(ctx, b.Pi.val.p, (int64_t)Pi.val.size(), b.Pi32);
      to32(ctx, b.PiT.val.p, (int64_t)PiT.val.size(), b.PiT32);
      b.slot_v = slot;
      slot += b.C;
      if (j - 1 == 0) {
        b.slot_t = slot++;
      } else {
        const int32_t* map_l = host_map(j - 1, hmap_l);
        HostCsr Pl = vector_nodal(m, g.coords, j - 1, map_l, map_0, nl, h->nv_nodes);
        HostCsr PlT = transpose(Pl);
        b.Cl = Pl.C;
        upload(ctx, Pl, b.Pil);
        upload(ctx, PlT, b.PilT);
        to32(ctx, b.Pil.val.p, (int64_t)Pl.val.size(), b.Pil32);
        to32(ctx, b.PilT.val.p, (int64_t)PlT.val.size(), b.PilT32);
        b.slot_t = slot;
        slot += b.Cl;
      }
    }
    h->blocks.push_back(std::move(b));
  }
  if (slot > MAXNV) return no("too many auxiliary copies for one cycle");
  h->nv = slot;
  alloc_slots(ctx, *h);
  h->zt64.alloc(ctx, std::max<int64_t>(h->N, 1));
  h->zt32.alloc(ctx, std::max<int64_t>(h->N, 1));
  smoother_diag(*h, s);
  // work estimate of one application in FP32: per level 3 + 2 sweeps passes
  // over the operator and the slot vectors, plus the transfers
  int64_t bytes = 0, launches = 4;
  for (size_t i = 0; i < h->L.size(); ++i) {
    const MgLevel& l = h->L[i];
    bytes += (int64_t)(2 * h->sweeps + 1) * (l.A.nnz * 8 + 12 * l.n * h->nv);
    launches += i + 1 < h->L.size() ? 3 + 2 * h->sweeps : 1;
  }
  for (const auto& b : h->blocks) {
    bytes += (b.D.nnz + b.DT.nnz) * 8 + (b.Pi.nnz + b.PiT.nnz) * 4 * (b.C + 1);
    launches += 6;
  }
  h->bytes32 = bytes;
  h->launches = launches;
  ctx->sync();
  return h.release();
}

void hx_refresh(Hx& h, Geom& g, const KktSystem& s) {
  Ctx* ctx = h.ctx;
  Csr A0;
  nodal_operator(g, h.bc, s, h.k, A0);
  MgLevel& l0 = h.L[0];
  FEEC_REQUIRE(A0.nnz == l0.A.nnz && A0.nrows == l0.n, FEEC_E_DEVICE, "nodal operator pattern changed");
  FEEC_HIP(hipMemcpyAsync(l0.A.val.p, A0.val.p, A0.nnz * 8, hipMemcpyDeviceToDevice, ctx->stream));
  level_refresh_derived(ctx, l0);
  hierarchy_values(h);
  smoother_diag(h, s);
  ctx->sync();
}

namespace {

template <class TP>
struct Sel;
template <>
struct Sel<double> {
  static const double* A(const MgLevel& l) { return l.A.val.p; }
  static const double* dinv(const MgLevel& l) { return l.dinv.p; }
  static const double* P(const MgLevel& l) { return l.P.val.p; }
  static const double* R(const MgLevel& l) { return l.R.val.p; }
  static double* r(MgLevel& l) { return l.r64.p; }
  static double* x(MgLevel& l) { return l.x64.p; }
  static double* y(MgLevel& l) { return l.y64.p; }
  static double* e(MgLevel& l) { return l.e64.p; }
  static const double* cinv(const Hx& h) { return h.cinv.p; }
  static const double* s(const Hx& h) { return h.s64.p; }
  static double* zt(Hx& h) { return h.zt64.p; }
  static const double* D(const HxBlock& b) { return b.D.val.p; }
  static const double* DT(const HxBlock& b) { return b.DT.val.p; }
  static const double* Pi(const HxBlock& b) { return b.Pi.val.p; }
  static const double* PiT(const HxBlock& b) { return b.PiT.val.p; }
  static const double* Pil(const HxBlock& b) { return b.Pil.val.p; }
  static const double* PilT(const HxBlock& b) { return b.PilT.val.p; }
  static const double* sl(const HxBlock& b) { return b.sl.p; }
  static double* t(HxBlock& b) { return b.t64.p; }
  static double* u(HxBlock& b) { return b.u64.p; }
};
template <>
struct Sel<float> {
  static const float* A(const MgLevel& l) { return l.A32.p; }
  static const float* dinv(const MgLevel& l) { return l.dinv32.p; }
  static const float* P(const MgLevel& l) { return l.P32.p; }
  static const float* R(const MgLevel& l) { return l.R32.p; }
  static float* r(MgLevel& l) { return l.r32.p; }
  static float* x(MgLevel& l) { return l.x32.p; }
  static float* y(MgLevel& l) { return l.y32.p; }
  static float* e(MgLevel& l) { return l.e32.p; }
  static const float* cinv(const Hx& h) { return h.cinv32.p; }
  static const float* s(const Hx& h) { return h.s32.p; }
  static float* zt(Hx& h) { return h.zt32.p; }
  static const float* D(const HxBlock& b) { return b.D32.p; }
  static const float* DT(const HxBlock& b) { return b.DT32.p; }
  static const float* Pi(const HxBlock& b) { return b.Pi32.p; }
  static const float* PiT(const HxBlock& b) { return b.PiT32.p; }
  static const float* Pil(const HxBlock& b) { return b.Pil32.p; }
  static const float* PilT(const HxBlock& b) { return b.PilT32.p; }
  static const float* sl(const HxBlock& b) { return b.sl32.p; }
  static float* t(HxBlock& b) { return b.t32.p; }
  static float* u(HxBlock& b) { return b.u32.p; }
};

// One V-cycle on level i: rhs in r(i), result in x(i) (iterate buffers swap).
template <class TP>
void vcycle(Enq& q, Hx& h, size_t i, const MinresS* S) {
  using X = Sel<TP>;
  MgLevel& l = h.L[i];
  const int nv = h.nv;
  if (i + 1 == h.L.size()) {
    q(k_mg_dense<TP>, (unsigned)std::max<int64_t>(l.n, 1), TPB, X::cinv(h), (const TP*)X::r(l), X::x(l),
      l.n, nv, S);
    return;
  }
  TP* x = X::x(l);
  TP* y = X::y(l);
  const TP* r = X::r(l);
  auto sweep = [&]() {
    q(k_mg_rowop<TP>, grid_for(l.n), TPB, l.A.row_ptr.p, l.A.col.p, X::A(l), X::dinv(l), r,
      (const TP*)x, y, l.n, nv, 0, S);
    std::swap(x, y);
  };
  q(k_mg_scale<TP>, grid_for(l.n), TPB, X::dinv(l), r, x, l.n, nv, S);
  for (int s = 1; s < h.sweeps; ++s) sweep();
  q(k_mg_rowop<TP>, grid_for(l.n), TPB, l.A.row_ptr.p, l.A.col.p, X::A(l), X::dinv(l), r,
    (const TP*)x, X::e(l), l.n, nv, 1, S);
  MgLevel& c = h.L[i + 1];
  q(k_mg_spmv<TP>, grid_for(c.n), TPB, l.R.row_ptr.p, l.R.col.p, X::R(l), (const TP*)X::e(l), X::r(c),
    c.n, l.n, nv, 0, S);
  vcycle<TP>(q, h, i + 1, S);
  q(k_mg_spmv<TP>, grid_for(l.n), TPB, l.P.row_ptr.p, l.P.col.p, X::P(l), (const TP*)X::x(c), x, l.n,
    c.n, nv, 1, S);
  for (int s = 0; s < h.sweeps; ++s) sweep();
  if (x != X::x(l))  // leave the result in x(l)
    q(k_hx_copy<TP, TP>, grid_for(l.n * nv), TPB, (const TP*)x, X::x(l), l.n * nv, 0, S);
}

template <class TIN, class TP, class TOUT>
void hx_apply(Enq& q, Hx& h, const TIN* r, TOUT* z, const MinresS* S, double* part) {
  using X = Sel<TP>;
  const int64_t n0 = h.nv_nodes;
  MgLevel& l0 = h.L[0];
  TP* slots = X::r(l0);
  TP* zt = X::zt(h);
  q(k_hx_init<TIN, TP>, grid_for(h.N), TPB, r, X::s(h), zt, h.n, h.N, S);
  for (HxBlock& b : h.blocks) {
    const TIN* rb = r + b.off;
    if (b.j == 0) {
      q(k_hx_copy<TIN, TP>, grid_for(b.n), TPB, rb, slots + b.slot_v * n0, b.n, 0, S);
      continue;
    }
    const int64_t nl = b.D.ncols;
    // t = D^T r
    q(k_hx_gather<TIN, TP>, grid_for(nl), TPB, b.DT.row_ptr.p, b.DT.col.p, X::DT(b), 1, rb, X::t(b), nl,
      S);
    if (b.j - 1 == 0)
      q(k_hx_copy<TP, TP>, grid_for(nl), TPB, (const TP*)X::t(b), slots + b.slot_t * n0, nl, 0, S);
    else
      q(k_hx_gather<TP, TP>, grid_for(n0), TPB, b.PilT.row_ptr.p, b.PilT.col.p, X::PilT(b), b.Cl,
        (const TP*)X::t(b), slots + b.slot_t * n0, n0, S);
    q(k_hx_gather<TIN, TP>, grid_for(n0), TPB, b.PiT.row_ptr.p, b.PiT.col.p, X::PiT(b), b.C, rb,
      slots + b.slot_v * n0, n0, S);
  }
  vcycle<TP>(q, h, 0, S);
  const TP* xs = X::x(l0);
  for (HxBlock& b : h.blocks) {
    TP* zb = zt + b.off;
    if (b.j == 0) {
      q(k_hx_copy<TP, TP>, grid_for(b.n), TPB, xs + b.slot_v * n0, zb, b.n, 1, S);
      continue;
    }
    const int64_t nl = b.D.ncols;
    // u = S_{j-1} t + (V t or Pi_{j-1} V Pi_{j-1}^T t)
    if (b.j - 1 == 0) {
      q(k_hx_scatter<TP>, grid_for(nl), TPB, (const int64_t*)nullptr, (const int32_t*)nullptr,
        (const TP*)nullptr, 0, (const TP*)nullptr, (int64_t)0, X::sl(b), (const TP*)X::t(b), X::u(b), nl,
        0, S);
      q(k_hx_copy<TP, TP>, grid_for(nl), TPB, xs + b.slot_t * n0, X::u(b), nl, 1, S);
    } else {
      q(k_hx_scatter<TP>, grid_for(nl), TPB, b.Pil.row_ptr.p, b.Pil.col.p, X::Pil(b), b.Cl,
        xs + b.slot_t * n0, n0, X::sl(b), (const TP*)X::t(b), X::u(b), nl, 0, S);
    }
    // z_j += D u + Pi V Pi^T r
    q(k_hx_scatter<TP>, grid_for(b.n), TPB, b.D.row_ptr.p, b.D.col.p, X::D(b), 1, (const TP*)X::u(b), nl,
      (const TP*)nullptr, (const TP*)nullptr, zb, b.n, 1, S);
    q(k_hx_scatter<TP>, grid_for(b.n), TPB, b.Pi.row_ptr.p, b.Pi.col.p, X::Pi(b), b.C,
      xs + b.slot_v * n0, n0, (const TP*)nullptr, (const TP*)nullptr, zb, b.n, 1, S);
  }
  q(k_hx_final<TIN, TP, TOUT>, red_grid(h.N), TPB, r, (const TP*)zt, z, h.N, S, part);
}

}  // namespace

void hx_enqueue(Enq& q, Hx& h, const double* r, double* z, const MinresS* S, double* part, bool fp32) {
  if (fp32)
    hx_apply<double, float, double>(q, h, r, z, S, part);
  else
    hx_apply<double, double, double>(q, h, r, z, S, part);
}

void hx_enqueue(Enq& q, Hx& h, const float* r, float* z, const MinresS* S, double* part) {
  hx_apply<float, float, float>(q, h, r, z, S, part);
}

}  // namespace feec
