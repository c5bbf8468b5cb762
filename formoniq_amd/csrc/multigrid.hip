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
// multi-right-hand-side cycle over "slots".
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
// Layout. Each level lives on its full vertex box ((g+1)^d vertices for g
// cells per axis). Its operator is stored by Kuhn-stencil diagonals: the
// neighbours of vertex i are i +- sum_{a in delta} stride_a for the 2^d - 1
// nonempty delta in {0,1}^d (every such diagonal of a box is an edge of the
// Kuhn triangulation, and the Galerkin product of the Kuhn-grid prolongation
// keeps that stencil), A[s * nb + i] for stencil point s (0 the diagonal,
// 1..2^d-1 the +delta, 2^d..2^{d+1}-2 the -delta), zero where the neighbour
// is not an unknown. Vertices that are not unknowns (the boundary of the
// relative complex, their coarse counterparts) are "dead": zero rows and
// columns and a zero smoother entry, so they stay zero through the cycle.
// Slot vectors are interleaved, x[i * nv + slot], and the cycle's kernels run
// one thread per (vertex, slot), so every vector access of a wavefront is
// contiguous and no column index is read. The prolongation and restriction
// are computed from grid coordinates, without matrices.
//
// Precision. The FP64 path runs every piece in FP64; the mixed paths in FP32
// (the FP64_PC32 solve rounds r to FP32 on entry and widens z on exit), where
// the Chebyshev steps multiply by an FP32 copy of P symmetrized exactly
// ((P_ij + P_ji)/2 rounded once for both), so the preconditioner stays a
// symmetric operator.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "enqueue.hpp"
#include "cell_geometry.hpp"
#include "feec_internal.hpp"

namespace feec {

namespace {

constexpr int TPB = 256;
constexpr int NBLK = 4096;
constexpr int MAXNV = 16;               // slots of one multi-right-hand-side cycle
constexpr int MAXC = 8;                 // values per entry of a transfer matrix
constexpr int64_t kCoarseMax = 200;     // largest level solved directly (unknowns)
constexpr int64_t kDenseMaxBox = 1024;  // largest coarsest box
constexpr double kOmega = 2.0 / 3.0;

inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + TPB - 1) / TPB); }
inline unsigned red_grid(int64_t n) {
  return (unsigned)std::min<int64_t>(NBLK, std::max<int64_t>(1, (n + TPB - 1) / TPB));
}

__device__ __forceinline__ bool skip(const MinresS* S) { return S && S->done; }

// The block's position in the index space, remapped so that each of the 8
// XCDs (to which blocks are dealt round-robin) works on one contiguous range:
// stencil neighbours and gathered vector entries then mostly hit that XCD's
// own L2 instead of being fetched once per XCD. A bijection on [0, gridDim.x).
__device__ __forceinline__ unsigned xcd_block() {
  const unsigned b = blockIdx.x, per = gridDim.x / 8;
  return b < per * 8 ? (b % 8) * per + b / 8 : b;
}

// A level's vertex box, passed by value.
struct Box {
  int g = 0;   // cells per axis
  int nb = 0;  // (g+1)^d vertices
  int st[4] = {0, 0, 0, 0};
};

// Linear offset of stencil point s on a box of strides st.
template <int D>
__device__ __host__ __forceinline__ int stencil_offset(int s, const int* st) {
  constexpr int H = (1 << D) - 1;
  const int mask = s <= H ? s : s - H;
  int o = 0;
#pragma unroll
  for (int a = 0; a < D; ++a)
    if (mask & (1 << a)) o += st[a];
  return s == 0 ? 0 : (s <= H ? o : -o);
}

// Stencil point of the coordinate difference `diff` (entries in {-1, 0, 1}),
// or -1 when it is not a Kuhn edge (mixed signs).
template <int D>
__device__ __host__ __forceinline__ int stencil_point(const int* diff) {
  int pos = 0, neg = 0;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    if (diff[a] == 1) pos |= 1 << a;
    if (diff[a] == -1) neg |= 1 << a;
  }
  if (pos && neg) return -1;
  return neg ? neg + (1 << D) - 1 : pos;
}

// ---------------------------------------------------------------------------
// the V-cycle on stencil levels: one thread per (vertex, slot)
// ---------------------------------------------------------------------------
// x = dinv r (the first sweep from x = 0)
template <class TP>
__global__ void __launch_bounds__(TPB) k_st_scale(const TP* __restrict__ dinv, const TP* __restrict__ r,
                                                  TP* __restrict__ x, unsigned nb, unsigned nv,
                                                  const MinresS* S) {
  if (skip(S)) return;
  const unsigned t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= nb * nv) return;
  x[t] = dinv[t / nv] * r[t];
}

// The sweep family, one thread per (vertex, slot). The operator is
// symmetric, so a row's -delta entries are read from the +delta diagonals at
// the neighbour (A(i, i - o) = A(i - o, i)): only the diagonal and the 2^d - 1
// +delta diagonals are fetched. MODE selects the iterate the sweep starts from:
//   0: out = x + dinv (r - A x)                  (a weighted Jacobi sweep)
//   1: out = r - A x, 0 on dead rows             (the residual)
//   2: the sweep from x = dinv r                 (the first two pre-sweeps in one)
// FINE marks the finest level's launches, which stream the operator from HBM,
// as their own instantiation (the coarse levels are cache-resident).
template <class TP, int MODE>
__device__ __forceinline__ TP st_iterate(int j, unsigned nv, unsigned v, const TP* __restrict__ x,
                                         const TP* __restrict__ dinv, const TP* __restrict__ r) {
  return MODE == 2 ? dinv[j] * r[(unsigned)j * nv + v] : x[(unsigned)j * nv + v];
}

template <class TP, int D, int MODE, bool FINE>
__global__ void __launch_bounds__(TPB) k_st_sweep(const TP* __restrict__ A, const TP* __restrict__ dinv,
                                                  const TP* __restrict__ r, const TP* __restrict__ x,
                                                  TP* __restrict__ out, Box b, unsigned nv, const MinresS* S) {
  constexpr int H = (1 << D) - 1;
  if (skip(S)) return;
  const unsigned t = xcd_block() * blockDim.x + threadIdx.x;
  const unsigned nb = (unsigned)b.nb;
  if (t >= nb * nv) return;
  const unsigned i = t / nv;
  const unsigned v = t - i * nv;
  const TP xi = st_iterate<TP, MODE>((int)i, nv, v, x, dinv, r);
  TP acc = A[i] * xi;
#pragma unroll
  for (int m = 1; m <= H; ++m) {
    const int o = stencil_offset<D>(m, b.st);
    // +delta: A(i, i + o) at row i; -delta: A(i, i - o) = A(i - o, i) at row i - o.
    // A neighbour outside the box has weight 0; its read is kept in range.
    int jp = (int)i + o, jm = (int)i - o;
    const TP ap = jp < (int)nb ? A[(size_t)m * nb + i] : TP(0);
    const TP am = jm >= 0 ? A[(size_t)m * nb + jm] : TP(0);
    jp = jp < (int)nb ? jp : (int)i;
    jm = jm >= 0 ? jm : (int)i;
    acc += ap * st_iterate<TP, MODE>(jp, nv, v, x, dinv, r);
    acc += am * st_iterate<TP, MODE>(jm, nv, v, x, dinv, r);
  }
  const TP d = dinv[i];
  const TP res = r[t] - acc;
  out[t] = MODE == 1 ? (d != TP(0) ? res : TP(0)) : xi + d * res;
}

// rc = P^T ef on the live coarse vertices (0 elsewhere): coarse vertex w
// collects its fine children 2w (weight 1) and 2w +- delta (weight 1/2).
template <class TP, int D>
__global__ void __launch_bounds__(TPB) k_st_restrict(const TP* __restrict__ ef, TP* __restrict__ rc,
                                                     const uint8_t* __restrict__ live_c, Box fb, Box cb,
                                                     unsigned nv, const MinresS* S) {
  constexpr int H = (1 << D) - 1;
  if (skip(S)) return;
  const unsigned t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= (unsigned)cb.nb * nv) return;
  const unsigned w = t / nv;
  const unsigned v = t - w * nv;
  if (!live_c[w]) {
    rc[t] = TP(0);
    return;
  }
  int wc[D];
  {
    unsigned r = w;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      wc[a] = (int)(r % (unsigned)(cb.g + 1));
      r /= (unsigned)(cb.g + 1);
    }
  }
  TP acc = TP(0);
#pragma unroll
  for (int s = 0; s <= 2 * H; ++s) {
    const int mask = s <= H ? s : s - H;
    const int sg = s <= H ? 1 : -1;
    bool in = true;
    int k = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const int c = 2 * wc[a] + ((mask >> a) & 1) * sg;
      in = in && c >= 0 && c <= fb.g;
      k += c * fb.st[a];
    }
    if (in) acc += (s == 0 ? TP(1) : TP(0.5)) * ef[(unsigned)k * nv + v];
  }
  rc[t] = acc;
}

// xf += P xc on the live fine vertices (dinv_f != 0).
template <class TP, int D>
__global__ void __launch_bounds__(TPB) k_st_prolong(const TP* __restrict__ xc, TP* __restrict__ xf,
                                                    const TP* __restrict__ dinv_f, Box fb, Box cb,
                                                    unsigned nv, const MinresS* S) {
  if (skip(S)) return;
  const unsigned t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= (unsigned)fb.nb * nv) return;
  const unsigned k = t / nv;
  const unsigned v = t - k * nv;
  if (dinv_f[k] == TP(0)) return;
  unsigned r = k;
  int lo = 0, hi = 0;
  bool odd = false;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const int c = (int)(r % (unsigned)(fb.g + 1));
    r /= (unsigned)(fb.g + 1);
    lo += (c >> 1) * cb.st[a];
    hi += ((c >> 1) + (c & 1)) * cb.st[a];
    odd = odd || (c & 1);
  }
  const TP add = odd ? TP(0.5) * (xc[(unsigned)lo * nv + v] + xc[(unsigned)hi * nv + v])
                     : xc[(unsigned)lo * nv + v];
  xf[t] += add;
}

// x = C r on the coarsest box, C dense (nb x nb, row-major, zero on dead
// rows and columns): one block per row, all slots.
template <class TP>
__global__ void __launch_bounds__(TPB) k_st_dense(const TP* __restrict__ C, const TP* __restrict__ r,
                                                  TP* __restrict__ x, int nb, int nv, const MinresS* S) {
  if (skip(S)) return;
  __shared__ TP sh[TPB / 64][MAXNV];
  const int i = blockIdx.x;
  TP acc[MAXNV];
#pragma unroll
  for (int v = 0; v < MAXNV; ++v) acc[v] = TP(0);
  for (int j = threadIdx.x; j < nb; j += blockDim.x) {
    const TP c = C[(size_t)i * nb + j];
#pragma unroll
    for (int v = 0; v < MAXNV; ++v)
      if (v < nv) acc[v] += c * r[j * nv + v];
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
    x[i * nv + threadIdx.x] = t;
  }
}

// The pattern side of P's FP32 copy, one thread per row i of the
// structurally symmetric P (sorted rows): for each entry (i, j) the position of
// (j, i) in row j (binary search), and the 16-bit offset j - i. flags[0] is
// set if a transpose entry is missing, flags[1] if an offset needs more than
// 16 bits.
__global__ void k_pc_pattern(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                             int64_t n, int64_t nnz, int32_t* __restrict__ tpos, int16_t* __restrict__ c16,
                             int* __restrict__ flags) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;  // one lane per entry
  if (p >= nnz) return;
  int64_t lo = 0, hi = n;  // the entry's row: the last i with rp[i] <= p
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if (rp[mid] <= p)
      lo = mid;
    else
      hi = mid;
  }
  const int64_t i = lo, j = col[p];
  lo = rp[j];
  hi = rp[j + 1];
  while (lo < hi) {  // first position in row j with column >= i
    const int64_t mid = (lo + hi) / 2;
    if (col[mid] < i)
      lo = mid + 1;
    else
      hi = mid;
  }
  const bool found = lo < rp[j + 1] && col[lo] == i;
  tpos[p] = found ? (int32_t)lo : 0;
  const int64_t off = j - i;
  const bool wide = off < -32767 || off > 32767;
  c16[p] = (int16_t)(wide ? 0 : off);
  if (!found) atomicOr(&flags[0], 1);
  if (wide) atomicOr(&flags[1], 1);
}

__global__ void k_zero_i32(int* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// operators of the hierarchy (FP64)
// ---------------------------------------------------------------------------
// out[map[q]] = val[src ? src[q] : q]
__global__ void k_scatter_map(const double* __restrict__ val, const int32_t* __restrict__ src,
                              const int32_t* __restrict__ map, int64_t n, double* __restrict__ out) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q < n) out[map[q]] = val[src ? src[q] : q];
}

// dinv = omega / A_ii on live rows, 0 on dead ones.
__global__ void k_st_dinv(const double* __restrict__ A, const uint8_t* __restrict__ live, int nb,
                          double omega, double* __restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) dinv[i] = live[i] && A[i] != 0.0 ? omega / A[i] : 0.0;
}

// A_c = P^T A_f P on the live coarse vertices, one thread per coarse vertex:
// for each fine child k of w (weight 1 or 1/2), each fine stencil neighbour
// l of k and each coarse parent p of l, the term w_k A_f(k, l) w_l lands on
// the coarse stencil point of p - w. bad[w] = 1 if a term falls outside the
// stencil (never, for the Kuhn prolongation; checked once at set-up).
template <int D>
__global__ void k_st_galerkin(const double* __restrict__ Af, const uint8_t* __restrict__ live_f,
                              const uint8_t* __restrict__ live_c, Box fb, Box cb, double* __restrict__ Ac,
                              uint8_t* __restrict__ bad) {
  constexpr int H = (1 << D) - 1;
  constexpr int NS = 2 * H + 1;
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= cb.nb) return;
  double acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = 0.0;
  uint8_t flag = 0;
  if (live_c[w]) {
    int wc[D];
    {
      int r = w;
      for (int a = 0; a < D; ++a) {
        wc[a] = r % (cb.g + 1);
        r /= (cb.g + 1);
      }
    }
    for (int cs = 0; cs < NS; ++cs) {  // fine child 2w + (+-delta)
      const int cm = cs <= H ? cs : cs - H;
      const int csg = cs <= H ? 1 : -1;
      int kc[D];
      bool in = true;
      int k = 0;
      for (int a = 0; a < D; ++a) {
        kc[a] = 2 * wc[a] + ((cm >> a) & 1) * csg;
        in = in && kc[a] >= 0 && kc[a] <= fb.g;
        k += kc[a] * fb.st[a];
      }
      if (!in || !live_f[k]) continue;
      const double wk = cs == 0 ? 1.0 : 0.5;
      for (int s = 0; s < NS; ++s) {  // fine neighbour l = k + (+-delta)
        const double a_kl = Af[(size_t)s * fb.nb + k];
        if (a_kl == 0.0) continue;
        const int sm = s <= H ? s : s - H;
        const int ssg = s <= H ? 1 : -1;
        int lo[D], hi[D];
        bool odd = false;
        for (int a = 0; a < D; ++a) {
          const int lc = kc[a] + ((sm >> a) & 1) * ssg;
          lo[a] = lc >> 1;
          hi[a] = (lc >> 1) + (lc & 1);
          odd = odd || (lc & 1);
        }
        for (int which = 0; which < (odd ? 2 : 1); ++which) {
          int pidx = 0, diff[D];
          for (int a = 0; a < D; ++a) {
            const int pc = which == 0 ? lo[a] : hi[a];
            pidx += pc * cb.st[a];
            diff[a] = pc - wc[a];
          }
          if (!live_c[pidx]) continue;
          const int slot = stencil_point<D>(diff);
          if (slot < 0) {
            flag = 1;
            continue;
          }
          const double term = wk * a_kl * (odd ? 0.5 : 1.0);
#pragma unroll
          for (int q = 0; q < NS; ++q)  // keeps the accumulators in registers
            if (q == slot) acc[q] += term;
        }
      }
    }
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) Ac[(size_t)s * cb.nb + w] = acc[s];
  bad[w] = flag;
}

// ---------------------------------------------------------------------------
// HX transfers between the KKT vector and the level-0 slots
// ---------------------------------------------------------------------------
// y[row * nv + slot0 + v] = sum_p val[p * C + v] x[col[p]] for v < C; LPR
// lanes share a row (entries strided over them, summed by shuffles).
template <class TX, class TP, int LPR>
__global__ void __launch_bounds__(TPB) k_hx_gather(const int64_t* __restrict__ rp,
                                                   const int32_t* __restrict__ col,
                                                   const TP* __restrict__ val, int C,
                                                   const TX* __restrict__ x, TP* __restrict__ y, int nv,
                                                   int slot0, int64_t nrows, const MinresS* S) {
  if (skip(S)) return;
  const int64_t gid = xcd_block() * (int64_t)blockDim.x + threadIdx.x;
  const int64_t row = gid / LPR;
  const int lane = (int)(gid % LPR);
  TP acc[MAXC];
#pragma unroll
  for (int v = 0; v < MAXC; ++v) acc[v] = TP(0);
  if (row < nrows) {
    for (int64_t p = rp[row] + lane; p < rp[row + 1]; p += LPR) {
      const TP xv = (TP)x[col[p]];
#pragma unroll
      for (int v = 0; v < MAXC; ++v)
        if (v < C) acc[v] += val[p * C + v] * xv;
    }
  }
  if (LPR > 1) {
#pragma unroll
    for (int v = 0; v < MAXC; ++v) {
      TP t = acc[v];
#pragma unroll
      for (int off = LPR / 2; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
      acc[v] = t;
    }
  }
  if (row < nrows && lane == 0)
    for (int v = 0; v < C; ++v) y[row * nv + slot0 + v] = acc[v];
}

// y[row] = (a ? a[row] b[row] : add ? y[row] : 0)
//          + sum_p sum_v val[p * C + v] X(col[p], v),
// X(c, v) = x0[c * nv + slot0 + v], except X(c, 0) = x0[..] + dl[c] r0[..]
// when `grad` (the grade-0 term S_0 t + V t of the gradient correction).
template <class TP, class TA, class TB>
__global__ void __launch_bounds__(TPB) k_hx_out(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                const TP* __restrict__ val, int C, int grad,
                                                const TP* __restrict__ x0, const TP* __restrict__ r0,
                                                const TP* __restrict__ dl, int nv, int slot0,
                                                const TA* __restrict__ a, const TB* __restrict__ b,
                                                TP* __restrict__ y, int64_t nrows, int add, const MinresS* S) {
  if (skip(S)) return;
  const int64_t i = xcd_block() * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  TP acc = a ? (TP)a[i] * (TP)b[i] : (add ? y[i] : TP(0));
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
    const int64_t base = (int64_t)col[p] * nv + slot0;
    TP x = x0[base];
    if (grad) x += dl[col[p]] * r0[base];
    acc += val[p * C] * x;
    for (int v = 1; v < C; ++v) acc += val[p * C + v] * x0[base + v];
  }
  y[i] = acc;
}

// The grade-1 block's scatter without a matrix: row e is the edge (v0, v1),
// v0 < v1 (box vertices), D_0 has -1 at v0 and +1 at v1 and Pi_1 the same D
// values eh = (x(v1) - x(v0)) / 2 at both ends:
//   z[e] = s[e] r[e] + (G(v1) - G(v0)) + sum_c eh[e, c] (X_c(v0) + X_c(v1)),
// G(v) = x0[v, slot0] + dl[v] r0[v, slot0] and X_c(v) = x0[v, slot0 + 1 + c].
// Vertices that are not unknowns hold zeros in every slot, so both ends are
// always read.
template <class TP, class TB, int D>
__global__ void __launch_bounds__(TPB) k_hx_out_edge(const int32_t* __restrict__ ev, const TP* __restrict__ eh,
                                                     const TP* __restrict__ x0, const TP* __restrict__ r0,
                                                     const TP* __restrict__ dl, int nv, int slot0,
                                                     const TP* __restrict__ s, const TB* __restrict__ r,
                                                     TP* __restrict__ z, int64_t nrows, const MinresS* S) {
  if (skip(S)) return;
  const int64_t e = xcd_block() * (int64_t)blockDim.x + threadIdx.x;
  if (e >= nrows) return;
  const int v0 = ev[2 * e], v1 = ev[2 * e + 1];
  const TP* a = x0 + (int64_t)v0 * nv + slot0;
  const TP* b = x0 + (int64_t)v1 * nv + slot0;
  TP acc = s[e] * (TP)r[e];
  acc += (b[0] + dl[v1] * r0[(int64_t)v1 * nv + slot0]) - (a[0] + dl[v0] * r0[(int64_t)v0 * nv + slot0]);
#pragma unroll
  for (int c = 0; c < D; ++c) acc += eh[e * D + c] * (a[1 + c] + b[1 + c]);
  z[e] = acc;
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

// z = zt on the preconditioned rows, z = r on the border rows, and the
// partial sums of r . z (red_grid(N) blocks).
template <class TIN, class TP, class TOUT>
__global__ void __launch_bounds__(TPB) k_hx_final(const TIN* __restrict__ r, const TP* __restrict__ zt,
                                                  TOUT* __restrict__ z, int64_t n, int64_t N, const MinresS* S,
                                                  double* __restrict__ part) {
  if (skip(S)) return;
  __shared__ double sh[TPB / 64];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    const TOUT v = i < n ? (TOUT)zt[i] : (TOUT)r[i];
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

__global__ void k_scale(const double* __restrict__ a, int64_t n, double s, double* __restrict__ b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = s * a[i];
}

// ---------------------------------------------------------------------------
// Chebyshev acceleration of the HX blocks: z = p(B P) B r per block
// ---------------------------------------------------------------------------
// coef: [1/theta_sigma, 1/theta_u, then per step k: rho_{k+1} rho_k,
// 2 rho_{k+1}/delta_sigma, 2 rho_{k+1}/delta_u]; row i < split is sigma's.

// y = P x, P the KKT preconditioner matrix (CSR); LPR lanes per row, each
// taking its entries U at a time as k_pc_spmv32 below does (the same sum).
template <class TV, class TP, int LPR, int U>
__global__ void __launch_bounds__(TPB) k_pc_spmv(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                 const TV* __restrict__ val, const TP* __restrict__ x,
                                                 TP* __restrict__ y, int64_t nrows, const MinresS* S) {
  if (skip(S)) return;
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t row = gid / LPR;
  const int lane = (int)(gid % LPR);
  TP acc = TP(0);
  if (row < nrows) {
    const int64_t e = rp[row + 1];
    int64_t p = rp[row] + lane;
    for (; p + (U - 1) * LPR < e; p += U * LPR) {
      int32_t c[U];
      TP v[U], xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        c[u] = col[p + u * LPR];
        v[u] = (TP)val[p + u * LPR];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) xv[u] = x[c[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) acc += v[u] * xv[u];
    }
    for (; p < e; p += LPR) acc += (TP)val[p] * x[col[p]];
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (row < nrows && lane == 0) y[row] = acc;
}

// P32[q] = FP32 of (P[q] + P[tpos[q]]) / 2, tpos[q] the entry of the
// transpose: an exactly symmetric single-precision copy of P. (Half precision
// is not enough: on the gradients, where the curl-curl part vanishes, P is
// its small mass part, which FP16 rounding of the large entries swamps.)
__global__ void k_sym_single(const double* __restrict__ val, const int32_t* __restrict__ tpos, int64_t n,
                             float* __restrict__ out) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q < n) out[q] = (float)(0.5 * (val[q] + val[tpos[q]]));
}

// y = P32 x, columns as int16 offsets from the row or int32; LPR lanes per row.
// Each lane takes its entries U at a time (loads, then gathers, then the
// products added in entry order): more loads in flight per lane and the same
// sum, bit for bit, as one entry at a time.
template <class TC, int LPR, int U>
__global__ void __launch_bounds__(TPB) k_pc_spmv32(const int64_t* __restrict__ rp, const TC* __restrict__ col,
                                                   const float* __restrict__ val, const float* __restrict__ x,
                                                   float* __restrict__ y, int64_t nrows, const MinresS* S) {
  if (skip(S)) return;
  const int64_t gid = xcd_block() * (int64_t)blockDim.x + threadIdx.x;
  const int64_t row = gid / LPR;
  const int lane = (int)(gid % LPR);
  const auto column = [row](TC c) { return std::is_same_v<TC, int16_t> ? row + (int64_t)c : (int64_t)c; };
  float acc = 0.f;
  if (row < nrows) {
    const int64_t e = rp[row + 1];
    int64_t p = rp[row] + lane;
    for (; p + (U - 1) * LPR < e; p += U * LPR) {
      TC c[U];
      float v[U], xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        c[u] = col[p + u * LPR];
        v[u] = val[p + u * LPR];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) xv[u] = x[column(c[u])];
#pragma unroll
      for (int u = 0; u < U; ++u) acc += v[u] * xv[u];
    }
    for (; p < e; p += LPR) acc += val[p] * x[column(col[p])];
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (row < nrows && lane == 0) y[row] = acc;
}

// res = zt, d = res / theta, z = 0
template <class TP>
__global__ void __launch_bounds__(TPB) k_hxc_first(const TP* __restrict__ zt, TP* __restrict__ res,
                                                   TP* __restrict__ d, TP* __restrict__ z, int64_t n,
                                                   int64_t split, const double* __restrict__ coef,
                                                   const MinresS* S) {
  if (skip(S)) return;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const TP r = zt[i];
  res[i] = r;
  d[i] = r * (TP)coef[i < split ? 0 : 1];
  z[i] = TP(0);
}

// step k: res -= zt, z += d, d = rho_{k+1} rho_k d + 2 rho_{k+1}/delta res
template <class TP>
__global__ void __launch_bounds__(TPB) k_hxc_step(const TP* __restrict__ zt, TP* __restrict__ res,
                                                  TP* __restrict__ d, TP* __restrict__ z, int64_t n,
                                                  int64_t split, const double* __restrict__ coef, int k,
                                                  const MinresS* S) {
  if (skip(S)) return;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* c = coef + 2 + 3 * k;
  const TP r = res[i] - zt[i];
  const TP di = d[i];
  res[i] = r;
  z[i] += di;
  d[i] = (TP)c[0] * di + (TP)c[i < split ? 1 : 2] * r;
}

// out = z + d on the preconditioned rows, r on the border rows, and the
// partial sums of r . out (red_grid(N) blocks).
template <class TIN, class TP, class TOUT>
__global__ void __launch_bounds__(TPB) k_hxc_final(const TIN* __restrict__ r, const TP* __restrict__ z,
                                                   const TP* __restrict__ d, TOUT* __restrict__ out, int64_t n,
                                                   int64_t N, const MinresS* S, double* __restrict__ part) {
  if (skip(S)) return;
  __shared__ double sh[TPB / 64];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    const TOUT v = i < n ? (TOUT)(z[i] + d[i]) : (TOUT)r[i];
    out[i] = v;
    acc += (double)r[i] * (double)v;
  }
  const double t = dev_block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// partial sums of x[i] y[i] over [lo, hi) (red_grid(hi - lo) blocks)
template <class TP>
__global__ void __launch_bounds__(TPB) k_dotr(const TP* __restrict__ x, const TP* __restrict__ y,
                                              int64_t lo, int64_t hi, double* __restrict__ part) {
  __shared__ double sh[TPB / 64];
  double acc = 0.0;
  for (int64_t i = lo + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < hi;
       i += (int64_t)gridDim.x * blockDim.x)
    acc += (double)x[i] * (double)y[i];
  const double t = dev_block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Lanczos state on the device (doubles): [0..1] rho, [2..3] rho0, [4..5] live,
// [6..7] alpha of the step, [8..9] beta of the step, [10..11] alphas kept,
// [12..13] betas kept, then alphas at 14 + 2 j + b and betas at
// 14 + 2 kMaxLanczos + 2 j + b.
constexpr int kMaxLanczos = 16;
// the sum of np partials, on a launch of one wavefront (every lane gets it)
__device__ __forceinline__ double lz_sum(const double* __restrict__ part, int np) {
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) v += part[i];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// rho = rho0 = r.Br per block; live when positive and finite
__global__ void k_lz_init(const double* __restrict__ part, int g0, int g1, double* __restrict__ lz) {
  for (int b = 0; b < 2; ++b) {
    const double v = lz_sum(part + b * NBLK, b ? g1 : g0);
    if (threadIdx.x == 0) {
      const bool has = (b ? g1 : g0) > 0;
      lz[b] = v;
      lz[2 + b] = v;
      lz[4 + b] = has && v > 0.0 && isfinite(v) ? 1.0 : 0.0;
      lz[10 + b] = 0.0;
      lz[12 + b] = 0.0;
    }
  }
}

// alpha = rho / p.Pp per live block (a block whose p.Pp is not positive stops)
__global__ void k_lz_alpha(const double* __restrict__ part, int g0, int g1, double* __restrict__ lz) {
  for (int b = 0; b < 2; ++b) {
    const double pw = lz_sum(part + b * NBLK, b ? g1 : g0);
    if (threadIdx.x == 0) {
      double a = 0.0;
      if (lz[4 + b] != 0.0) {
        if (!(pw > 0.0) || !isfinite(pw)) {
          lz[4 + b] = 0.0;
        } else {
          a = lz[b] / pw;
          const int j = (int)lz[10 + b];
          lz[14 + 2 * j + b] = a;
          lz[10 + b] = j + 1;
        }
      }
      lz[6 + b] = a;
    }
  }
}

// beta = r.Br (new) / rho per live block (a block whose r.Br fell below stop
// of its start stops), rho = r.Br
__global__ void k_lz_beta(const double* __restrict__ part, int g0, int g1, double stop, double* __restrict__ lz) {
  for (int b = 0; b < 2; ++b) {
    const double rn = lz_sum(part + b * NBLK, b ? g1 : g0);
    if (threadIdx.x == 0) {
      double bt = 0.0;
      if (lz[4 + b] != 0.0) {
        if (!(rn > stop * lz[2 + b]) || !isfinite(rn)) {
          lz[4 + b] = 0.0;
        } else {
          bt = rn / lz[b];
          const int j = (int)lz[12 + b];
          lz[14 + 2 * kMaxLanczos + 2 * j + b] = bt;
          lz[12 + b] = j + 1;
          lz[b] = rn;
        }
      }
      lz[8 + b] = bt;
    }
  }
}

// y -= alpha_b x and p = z + beta_b p with the step's scalars from lz
template <class TP>
__global__ void k_lz_axpy(const TP* __restrict__ x, TP* __restrict__ y, int64_t n, int64_t split,
                          const double* __restrict__ lz) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] -= (TP)lz[6 + (i < split ? 0 : 1)] * x[i];
}
template <class TP>
__global__ void k_lz_xpby(const TP* __restrict__ z, TP* __restrict__ p, int64_t n, int64_t split,
                          const double* __restrict__ lz) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + (TP)lz[8 + (i < split ? 0 : 1)] * p[i];
}

// y += a_b x, a_0 on [0, split) and a_1 on [split, n)
template <class TP>
__global__ void k_block_axpy(const TP* __restrict__ x, TP* __restrict__ y, int64_t n, int64_t split, double a0,
                             double a1) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] += (TP)(i < split ? a0 : a1) * x[i];
}

// p = z + b_b p
template <class TP>
__global__ void k_block_xpby(const TP* __restrict__ z, TP* __restrict__ p, int64_t n, int64_t split, double b0,
                             double b1) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + (TP)(i < split ? b0 : b1) * p[i];
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

}  // namespace

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
  Box box;
  int64_t live = 0;     // unknowns
  DBuf<uint8_t> alive;  // per box vertex
  DBuf<double> A;       // stencil diagonals, NS * nb
  DBuf<float> A32;
  DBuf<double> dinv;  // omega / diag on live rows, 0 on dead ones
  DBuf<float> dinv32;
  DBuf<uint8_t> bad;  // Galerkin check (coarse levels)
  // slot vectors (nb * nv): rhs, two iterates, residual
  DBuf<double> r64, x64, y64, e64;
  DBuf<float> r32, x32, y32, e32;
};

// A CSR transfer matrix with C values per entry, FP64 and FP32.
struct XCsr {
  Csr m;
  DBuf<float> v32;
  int C = 1;
  double avg_row = 0.0;
};

struct HxBlock {
  int j = -1;  // grade of the block
  int64_t off = 0, n = 0;
  // into the level-0 slots from the block's residual: j = 0 the nodal DOF
  // map, j = 1 [D_0^T | Pi_1^T] (the gradient slot, then the Pi slots), j = 2 Pi_2^T
  XCsr g_main;
  int slot_main = 0;
  // back onto the block: j = 0 the map, j = 1 [D_0 | Pi_1] (gradient), j = 2 Pi_2
  XCsr o_main;
  // j = 2 only: t = D_1^T r, Pi_1^T t into the slots, u = S_1 t + Pi_1 (V ..), z += D_1 u
  XCsr g_t, g_tl, o_ul, o_d;
  int slot_tl = 0;
  DBuf<double> sl, t64, u64;  // sl = S_1 on the grade-1 DOFs
  DBuf<float> sl32, t32, u32;
  // j = 1: the scatter without a matrix (k_hx_out_edge) when D_0's signs follow
  // the sorted ends: the ends per row and the Pi_1 values
  bool edge_free = false;
  DBuf<int32_t> ev;
  DBuf<double> eh64;
  DBuf<float> eh32;
};

struct Hx {
  Ctx* ctx = nullptr;
  int d = 0, k = 0, bc = 0, grid = 0, sweeps = 2;
  int64_t n = 0, N = 0;
  int nv = 0;              // slots
  std::vector<MgLevel> L;  // finest first; the last is solved densely
  // level-0 operator values: entry q lands on stencil position l0_map[q] and
  // comes from P's value l0_src[q] (k <= 1: the nodal operator is P's block)
  // or from the q-th value of the assembled nodal operator (k = 2)
  DBuf<int32_t> l0_map, l0_src;
  int64_t l0_nnz = 0;
  bool l0_from_p = false;
  DBuf<double> cinv;
  DBuf<float> cinv32;
  std::vector<HxBlock> blocks;
  DBuf<double> s64, zt64;  // KKT-vector smoother diagonal and work vector
  DBuf<float> s32, zt32;
  // Chebyshev acceleration (cheb_m applications of B per block application)
  int cheb_m = 1;
  double frac = 0.25;
  bool spectrum_fp32 = false;  // estimate lambda_max with the FP32 cycle (mixed-precision problems)
  double lam[2] = {0.0, 0.0};  // lambda_max of B P per block (sigma, u)
  int64_t split = 0;
  const int64_t* P_rp = nullptr;
  const int32_t* P_col = nullptr;
  const double* P_val = nullptr;
  int64_t P_nnz = 0;
  DBuf<float> P32;              // exactly symmetric FP32 copy: the FP32 Chebyshev steps' P
  DBuf<int32_t> P_tpos;         // entry of P's transpose, per entry
  DBuf<int16_t> P_col16;        // column offsets from the row, when all fit 16 bits
  DBuf<double> coef, pvec, part, lz;
  DBuf<double> res64, d64, zc64, w64;
  DBuf<float> res32, d32, zc32, w32;
  int64_t bytes32 = 0, launches = 0;
  // the spectrum estimate's launches, captured once and replayed per refresh
  bool use_graph = true;         // replay the estimate from a graph (else launch one by one)
  bool lz_direct_ready = false;  // the start vector is on the device
  hipGraph_t lz_graph = nullptr;
  hipGraphExec_t lz_exec = nullptr;
  std::vector<std::shared_ptr<void>> lz_keep;
  ~Hx() {
    if (lz_exec) (void)hipGraphExecDestroy(lz_exec);
    if (lz_graph) (void)hipGraphDestroy(lz_graph);
  }
};

void hx_free(Hx* h) { delete h; }
int64_t hx_bytes_per_apply(const Hx& h) { return h.bytes32; }
int hx_launches_per_apply(const Hx& h) { return (int)h.launches; }

namespace {

void upload(Ctx* ctx, const HostCsr& h, XCsr& out) {
  Csr& c = out.m;
  c.ctx = ctx;
  c.nrows = h.nrows;
  c.ncols = h.ncols;
  c.nnz = (int64_t)h.col.size();
  c.row_ptr.alloc(ctx, h.nrows + 1);
  c.col.alloc(ctx, std::max<int64_t>(c.nnz, 1));
  c.val.alloc(ctx, std::max<size_t>(h.val.size(), 1));
  copy_to_device(ctx, c.row_ptr.p, h.rp.data(), (h.nrows + 1) * 8);
  if (c.nnz) copy_to_device(ctx, c.col.p, h.col.data(), c.nnz * 4);
  if (!h.val.empty()) copy_to_device(ctx, c.val.p, h.val.data(), h.val.size() * 8);
  out.C = h.C;
  out.avg_row = h.nrows ? (double)c.nnz / (double)h.nrows : 0.0;
  out.v32.alloc(ctx, std::max<size_t>(h.val.size(), 1));
  if (!h.val.empty())
    k_convert<double, float><<<grid_for((int64_t)h.val.size()), TPB, 0, ctx->stream>>>(
        c.val.p, out.v32.p, (int64_t)h.val.size());
  FEEC_HIP(hipGetLastError());
}

void to32(Ctx* ctx, const double* a, int64_t n, DBuf<float>& out) {
  if ((int64_t)out.n != std::max<int64_t>(n, 1)) out.alloc(ctx, std::max<int64_t>(n, 1));
  if (n) k_convert<double, float><<<grid_for(n), TPB, 0, ctx->stream>>>(a, out.p, n);
  FEEC_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// transfer operators of the auxiliary spaces, built on the device
// ---------------------------------------------------------------------------
// The ascending vertices of the grade-J simplex s (J >= 1), decoded from its
// cell slot rep[s] = cell * C(d+1, J+1) + local face (faces in mask order).
template <int J>
__device__ inline void simplex_verts(const int32_t* cells, const int64_t* rep, int d, int64_t s,
                                     int32_t (&v)[J + 1]) {
  const int nv = d + 1;
  if (J == d) {
    for (int p = 0; p <= J; ++p) v[p] = cells[s * nv + p];
    return;
  }
  const int nf = binom(nv, J + 1);
  const int64_t slot = rep[s];
  const unsigned msk = nth_subset(nv, J + 1, (int)(slot % nf));
  const int64_t cell = slot / nf;
  for (int p = 0; p <= J; ++p) v[p] = cells[cell * nv + nth_elem(msk, p)];
}

__global__ void k_fill_i32(int32_t* a, int64_t n, int32_t v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// order[map[s]] = s: the simplex behind each block row (map < 0: no row)
__global__ void k_invert_map(const int32_t* map, int64_t ns, int64_t nrows, int32_t* order) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const int32_t r = map[s];
  if (r >= 0 && r < nrows) order[r] = (int32_t)s;
}

// the box column of each vertex: itself when it carries a DOF, else -1
__global__ void k_live_col(const int32_t* dof0, int64_t nv, int32_t* live) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v < nv) live[v] = dof0[v] >= 0 ? (int32_t)v : -1;
}

// entries of each vector-nodal row (its vertices that keep a column); the
// extra entry at nrows is 0 so that the exclusive scan ends on the total
template <int J>
__global__ void k_vnodal_count(const int32_t* cells, const int64_t* rep, int d, const int32_t* order,
                               const int32_t* colmap, int64_t nrows, unsigned long long* cnt) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r > nrows) return;
  const int64_t s = r == nrows ? -1 : (order ? order[r] : r);
  unsigned long long c = 0;
  if (s >= 0) {
    int32_t v[J + 1];
    simplex_verts<J>(cells, rep, d, s, v);
    for (int p = 0; p <= J; ++p) c += (colmap ? colmap[v[p]] : v[p]) >= 0;
  }
  cnt[r] = c;
}

// Vector-nodal interpolation of grade J (formoniq/src/hx.rs:86-112): the row
// of the J-simplex sigma has an entry per vertex of sigma with C = C(d, J)
// values, value I = (1/(J+1)!) det(E_I), E the edge vectors of sigma from its
// first vertex and E_I their components I (colex J-subsets of the d axes);
// entries sorted by column.
template <int J>
__global__ void k_vnodal_fill(const int32_t* cells, const int64_t* rep, int d, const int32_t* order,
                              const int32_t* colmap, int64_t nrows, const double* X, double factor,
                              const int64_t* rp, int32_t* col, double* val) {
#pragma clang fp contract(off)
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const int64_t s = order ? order[r] : r;
  if (s < 0) return;
  int32_t v[J + 1];
  simplex_verts<J>(cells, rep, d, s, v);
  double E[J][4];
  for (int i = 0; i < J; ++i)
    for (int a = 0; a < d; ++a) E[i][a] = X[(int64_t)v[i + 1] * d + a] - X[(int64_t)v[0] * d + a];
  double w[6];
  int C = 0;
  for (unsigned mask = 1; mask < (1u << d); ++mask) {  // colex order of subsets = increasing mask
    if (popc(mask) != J) continue;
    const int a0 = nth_elem(mask, 0);
    double det;
    if constexpr (J == 1) {
      det = E[0][a0];
    } else {
      const int a1 = nth_elem(mask, 1);
      det = E[0][a0] * E[1][a1] - E[0][a1] * E[1][a0];
    }
    w[C++] = factor * det;
  }
  int32_t ec[J + 1];
  int ne = 0;
  for (int p = 0; p <= J; ++p) {
    const int32_t c = colmap ? colmap[v[p]] : v[p];
    if (c < 0) continue;
    int q = ne++;
    while (q > 0 && ec[q - 1] > c) {
      ec[q] = ec[q - 1];
      --q;
    }
    ec[q] = c;
  }
  const int64_t p0 = rp[r];
  for (int e = 0; e < ne; ++e) {
    col[p0 + e] = ec[e];
    for (int I = 0; I < C; ++I) val[(p0 + e) * C + I] = w[I];
  }
}

// COO of the transpose: entry p of row r becomes (col[p], r) from source p
__global__ void k_transpose_coo(const int64_t* rp, const int32_t* col, int64_t nrows, int32_t* orow,
                                int32_t* ocol, uint32_t* osrc) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  for (int64_t p = rp[r]; p < rp[r + 1]; ++p) {
    orow[p] = col[p];
    ocol[p] = (int32_t)r;
    osrc[p] = (uint32_t)p;
  }
}

// the C values of each (duplicate-free) plan entry from its one source
__global__ void k_gather_entries(const int64_t* seg, const uint32_t* src, const double* in, int C,
                                 int64_t nnz, double* out) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= nnz * C) return;
  const int64_t e = q / C;
  out[q] = in[(int64_t)src[seg[e]] * C + (q - e * C)];
}

// [a | b] on one pattern: a's C_a values, then b's C_b values, per entry;
// bad[0] set when the patterns differ
__global__ void k_interleave(const int64_t* arp, const int32_t* acol, const double* av, int ca,
                             const int64_t* brp, const int32_t* bcol, const double* bv, int cb,
                             int64_t nrows, double* ov, int* bad) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  bool same = arp[r] == brp[r] && arp[r + 1] == brp[r + 1];
  const int co = ca + cb;
  for (int64_t p = arp[r]; p < arp[r + 1]; ++p) {
    same = same && acol[p] == bcol[p];
    for (int v = 0; v < ca; ++v) ov[p * co + v] = av[p * ca + v];
    for (int v = 0; v < cb; ++v) ov[p * co + ca + v] = bv[p * cb + v];
  }
  if (!same) atomicOr(bad, 1);
}

// The matrix-free edge scatter: the two ends of each edge row and Pi_1's row
// values (1/2!) (x(v1) - x(v0)); bad[0] set unless every D_0 entry of the row
// is -1 at the smaller end or +1 at the larger one.
__global__ void k_edge_ends(const int32_t* cells, const int64_t* rep, int d, const int32_t* order,
                            int64_t nrows, const int64_t* drp, const int32_t* dcol, const double* dval,
                            const double* X, int32_t* ev, double* eh, int* bad) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const int64_t s = order ? order[r] : r;
  if (s < 0) {
    atomicOr(bad, 1);
    return;
  }
  int32_t v[2];
  simplex_verts<1>(cells, rep, d, s, v);
  bool ok = v[0] < v[1];
  for (int64_t p = drp[r]; p < drp[r + 1]; ++p)
    ok = ok && ((dcol[p] == v[0] && dval[p] == -1.0) || (dcol[p] == v[1] && dval[p] == 1.0));
  if (!ok) atomicOr(bad, 1);
  ev[2 * r] = v[0];
  ev[2 * r + 1] = v[1];
  for (int c = 0; c < d; ++c)
    eh[r * d + c] = (1.0 / 2.0) * (X[(int64_t)v[1] * d + c] - X[(int64_t)v[0] * d + c]);
}

// avg_row and the FP32 copy of a finished transfer operator
void finish_x(Ctx* ctx, XCsr& out) {
  Csr& c = out.m;
  if (!c.col.p) c.col.alloc(ctx, 1);
  if (!c.val.p) c.val.alloc(ctx, 1);
  const int64_t nv = c.nnz * out.C;
  out.avg_row = c.nrows ? (double)c.nnz / (double)c.nrows : 0.0;
  to32(ctx, c.val.p, nv, out.v32);
}

// the simplex behind each of the nrows block rows of grade j (null: identity)
void row_order(Mesh& m, const int32_t* map, int j, int64_t nrows, DBuf<int32_t>& order) {
  Ctx* ctx = m.ctx;
  order.reset();
  if (!map) return;
  order.alloc(ctx, std::max<int64_t>(nrows, 1));
  if (nrows) k_fill_i32<<<grid_for(nrows), TPB, 0, ctx->stream>>>(order.p, nrows, -1);
  if (m.n[j]) k_invert_map<<<grid_for(m.n[j]), TPB, 0, ctx->stream>>>(map, m.n[j], nrows, order.p);
  FEEC_HIP(hipGetLastError());
}

// Pi_j on the block rows `order` (null: all j-simplices), columns colmap (null: the vertices).
void vector_nodal_dev(Mesh& m, const double* X, int j, const int32_t* order, const int32_t* colmap,
                      int64_t nrows, int64_t ncols, XCsr& out) {
  Ctx* ctx = m.ctx;
  const int d = m.d;
  FEEC_REQUIRE(j == 1 || j == 2, FEEC_E_DEVICE, "vector-nodal interpolation of an unsupported grade");
  const int C = binom(d, j);
  Csr& c = out.m;
  c.ctx = ctx;
  c.nrows = nrows;
  c.ncols = ncols;
  c.row_ptr.alloc(ctx, nrows + 1);
  DBuf<unsigned long long> cnt(ctx, nrows + 1);
  const int32_t* cells = m.faces[0].p;
  const int64_t* rep = m.rep[j].p;
  const unsigned gc = grid_for(nrows + 1), gf = grid_for(std::max<int64_t>(nrows, 1));
  if (j == 1)
    k_vnodal_count<1><<<gc, TPB, 0, ctx->stream>>>(cells, rep, d, order, colmap, nrows, cnt.p);
  else
    k_vnodal_count<2><<<gc, TPB, 0, ctx->stream>>>(cells, rep, d, order, colmap, nrows, cnt.p);
  exclusive_scan_counts(ctx, cnt.p, c.row_ptr.p, nrows + 1);
  copy_to_host(ctx, &c.nnz, c.row_ptr.p + nrows, 8);
  c.col.alloc(ctx, std::max<int64_t>(c.nnz, 1));
  c.val.alloc(ctx, std::max<int64_t>(c.nnz * C, 1));
  double fact = 1.0;
  for (int i = 2; i <= j + 1; ++i) fact *= i;
  if (nrows) {
    if (j == 1)
      k_vnodal_fill<1><<<gf, TPB, 0, ctx->stream>>>(cells, rep, d, order, colmap, nrows, X, 1.0 / fact,
                                                    c.row_ptr.p, c.col.p, c.val.p);
    else
      k_vnodal_fill<2><<<gf, TPB, 0, ctx->stream>>>(cells, rep, d, order, colmap, nrows, X, 1.0 / fact,
                                                    c.row_ptr.p, c.col.p, c.val.p);
  }
  FEEC_HIP(hipGetLastError());
  out.C = C;
  finish_x(ctx, out);
}

// the transpose of a transfer operator, C values per entry carried along
void transpose_dev(Ctx* ctx, const XCsr& a, XCsr& out) {
  const Csr& m = a.m;
  const int64_t nnz = m.nnz;
  DBuf<int32_t> rows(ctx, std::max<int64_t>(nnz, 1)), cols(ctx, std::max<int64_t>(nnz, 1));
  DBuf<uint32_t> srcs(ctx, std::max<int64_t>(nnz, 1));
  if (nnz)
    k_transpose_coo<<<grid_for(m.nrows), TPB, 0, ctx->stream>>>(m.row_ptr.p, m.col.p, m.nrows, rows.p,
                                                                 cols.p, srcs.p);
  CooPlan plan;
  coo_plan_build(ctx, plan, m.ncols, m.nrows, nnz, rows.p, cols.p, srcs.p);
  FEEC_REQUIRE(plan.nnz == nnz, FEEC_E_DEVICE, "a transfer operator with repeated entries");
  Csr& c = out.m;
  c.ctx = ctx;
  c.nrows = m.ncols;
  c.ncols = m.nrows;
  c.nnz = nnz;
  c.row_ptr = std::move(plan.row_ptr);
  c.col = std::move(plan.col);
  c.val.alloc(ctx, std::max<int64_t>(nnz * a.C, 1));
  if (nnz)
    k_gather_entries<<<grid_for(nnz * a.C), TPB, 0, ctx->stream>>>(plan.seg.p, plan.src.p, m.val.p, a.C,
                                                                    nnz, c.val.p);
  FEEC_HIP(hipGetLastError());
  out.C = a.C;
  finish_x(ctx, out);
}

// The coboundary D_{j-1} on the block's DOFs: rows the j-DOFs (rowmap: the
// DOF of each j-simplex, null: identity), columns mapped by colmap (-1:
// dropped), entries sorted by column.
void coboundary_dev(Mesh& m, int j, const int32_t* rowmap, int64_t nrows, const int32_t* colmap,
                    int64_t ncols, XCsr& out) {
  Csr full;
  mesh_coboundary(m, j - 1, full);
  csr_select(full, rowmap, nrows, colmap, ncols, out.m);
  out.C = 1;
  finish_x(m.ctx, out);
}

// [a | b]: the entries of two operators of one pattern, a's values first;
// raises bad[0] on the device when the patterns differ.
void interleave_dev(Ctx* ctx, const XCsr& a, const XCsr& b, XCsr& out, int* bad) {
  FEEC_REQUIRE(a.m.nrows == b.m.nrows && a.m.nnz == b.m.nnz, FEEC_E_DEVICE, "transfer patterns differ");
  const Csr& am = a.m;
  Csr& c = out.m;
  c.ctx = ctx;
  c.nrows = am.nrows;
  c.ncols = am.ncols;
  c.nnz = am.nnz;
  out.C = a.C + b.C;
  c.row_ptr.alloc(ctx, am.nrows + 1);
  c.col.alloc(ctx, std::max<int64_t>(am.nnz, 1));
  c.val.alloc(ctx, std::max<int64_t>(am.nnz * out.C, 1));
  FEEC_HIP(hipMemcpyAsync(c.row_ptr.p, am.row_ptr.p, (am.nrows + 1) * 8, hipMemcpyDeviceToDevice,
                          ctx->stream));
  if (am.nnz) {
    FEEC_HIP(hipMemcpyAsync(c.col.p, am.col.p, am.nnz * 4, hipMemcpyDeviceToDevice, ctx->stream));
    k_interleave<<<grid_for(am.nrows), TPB, 0, ctx->stream>>>(am.row_ptr.p, am.col.p, am.val.p, a.C,
                                                               b.m.row_ptr.p, b.m.col.p, b.m.val.p, b.C,
                                                               am.nrows, c.val.p, bad);
  }
  FEEC_HIP(hipGetLastError());
  finish_x(ctx, out);
}

// The nodal operator hdif_gram(0) on the complex's grade-0 DOFs: P's block
// when the KKT preconditioner holds it (k <= 1), else assembled on its own.
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

int nstencil(int d) { return (1 << (d + 1)) - 1; }

int host_stencil_offset(int d, int s, const int* st) {
  switch (d) {
    case 1: return stencil_offset<1>(s, st);
    case 2: return stencil_offset<2>(s, st);
    case 3: return stencil_offset<3>(s, st);
    default: return stencil_offset<4>(s, st);
  }
}

// Entries of the nodal block of rows and columns [lo, hi) of a CSR, per
// row; the extra entry at hi - lo is 0 so that the exclusive scan ends on the total.
__global__ void k_l0_count(const int64_t* rp, const int32_t* col, int64_t lo, int64_t hi,
                           unsigned long long* cnt) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t n = hi - lo;
  if (i > n) return;
  unsigned long long c = 0;
  if (i < n)
    for (int64_t p = rp[lo + i]; p < rp[lo + i + 1]; ++p) c += col[p] >= lo && col[p] < hi;
  cnt[i] = c;
}

// The level-0 stencil slot (point * nb + row vertex) of each entry of the
// nodal block, in row order, with its source entry when src is given;
// bad[0] set for an entry off the Kuhn stencil. box_of: the vertex of each
// nodal DOF (null: the DOF is the vertex).
template <int D>
__global__ void k_l0_fill(const int64_t* rp, const int32_t* col, int64_t lo, int64_t hi,
                          const int32_t* box_of, Box b0, const int64_t* start, int32_t* map,
                          int32_t* src, int* bad) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= hi - lo) return;
  const int g1 = b0.g + 1;
  const int bi = box_of ? box_of[i] : (int)i;
  int64_t q = start[i];
  bool ok = bi >= 0;
  for (int64_t p = rp[lo + i]; p < rp[lo + i + 1]; ++p) {
    const int32_t c = col[p];
    if (c < lo || c >= hi) continue;
    int ri = bi, rj = box_of ? box_of[c - lo] : (int)(c - lo);
    ok = ok && rj >= 0;
    int diff[D];
    for (int a = 0; a < D; ++a) {
      diff[a] = rj % g1 - ri % g1;
      ok = ok && diff[a] >= -1 && diff[a] <= 1;
      ri /= g1;
      rj /= g1;
    }
    const int sp = stencil_point<D>(diff);
    ok = ok && sp >= 0;
    map[q] = ok ? (int32_t)((int64_t)sp * b0.nb + bi) : 0;
    if (src) src[q] = (int32_t)p;
    ++q;
  }
  if (!ok) atomicOr(bad, 1);
}

void launch_galerkin(Ctx* ctx, int d, MgLevel& f, MgLevel& c) {
  const unsigned gr = grid_for(c.box.nb);
  switch (d) {
    case 1: k_st_galerkin<1><<<gr, TPB, 0, ctx->stream>>>(f.A.p, f.alive.p, c.alive.p, f.box, c.box, c.A.p, c.bad.p); break;
    case 2: k_st_galerkin<2><<<gr, TPB, 0, ctx->stream>>>(f.A.p, f.alive.p, c.alive.p, f.box, c.box, c.A.p, c.bad.p); break;
    case 3: k_st_galerkin<3><<<gr, TPB, 0, ctx->stream>>>(f.A.p, f.alive.p, c.alive.p, f.box, c.box, c.A.p, c.bad.p); break;
    default: k_st_galerkin<4><<<gr, TPB, 0, ctx->stream>>>(f.A.p, f.alive.p, c.alive.p, f.box, c.box, c.A.p, c.bad.p); break;
  }
  FEEC_HIP(hipGetLastError());
}

void level_derived(Ctx* ctx, int d, MgLevel& l) {
  const int64_t ns = (int64_t)nstencil(d) * l.box.nb;
  k_st_dinv<<<grid_for(l.box.nb), TPB, 0, ctx->stream>>>(l.A.p, l.alive.p, l.box.nb, kOmega, l.dinv.p);
  FEEC_HIP(hipGetLastError());
  to32(ctx, l.dinv.p, l.box.nb, l.dinv32);
  to32(ctx, l.A.p, ns, l.A32);
}

// Level operators from the finest one's values: the level-0 stencil from the
// nodal CSR, Galerkin products down the hierarchy, the coarsest inverse.
// Buffers stay where they are, so captured graphs stay valid.
void hierarchy_values(Hx& h, const double* vals, bool check) {
  Ctx* ctx = h.ctx;
  const int d = h.d;
  MgLevel& l0 = h.L[0];
  FEEC_HIP(hipMemsetAsync(l0.A.p, 0, l0.A.n * sizeof(double), ctx->stream));
  if (h.l0_nnz)
    k_scatter_map<<<grid_for(h.l0_nnz), TPB, 0, ctx->stream>>>(vals, h.l0_from_p ? h.l0_src.p : nullptr,
                                                                h.l0_map.p, h.l0_nnz, l0.A.p);
  FEEC_HIP(hipGetLastError());
  level_derived(ctx, d, l0);
  for (size_t i = 0; i + 1 < h.L.size(); ++i) {
    MgLevel& f = h.L[i];
    MgLevel& c = h.L[i + 1];
    launch_galerkin(ctx, d, f, c);
    level_derived(ctx, d, c);
    if (check) {
      std::vector<uint8_t> bad(c.box.nb);
      copy_to_host(ctx, bad.data(), c.bad.p, c.box.nb);
      for (uint8_t b : bad) FEEC_REQUIRE(!b, FEEC_E_DEVICE, "Galerkin product left the Kuhn stencil");
    }
  }
  // the coarsest level: dense inverse on its live vertices, embedded in its box
  MgLevel& last = h.L.back();
  const int nb = last.box.nb, NS = nstencil(d);
  std::vector<double> A((size_t)NS * nb);
  std::vector<uint8_t> alive(nb);
  copy_to_host(ctx, A.data(), last.A.p, A.size() * 8);
  copy_to_host(ctx, alive.data(), last.alive.p, nb);
  std::vector<int> pos(nb, -1), live;
  for (int i = 0; i < nb; ++i)
    if (alive[i]) {
      pos[i] = (int)live.size();
      live.push_back(i);
    }
  const int64_t nl = (int64_t)live.size();
  std::vector<double> dense(nl * nl, 0.0);
  for (int64_t r = 0; r < nl; ++r) {
    const int i = live[r];
    for (int s = 0; s < NS; ++s) {
      const double a = A[(size_t)s * nb + i];
      if (a == 0.0) continue;
      const int jv = i + host_stencil_offset(d, s, last.box.st);
      FEEC_REQUIRE(jv >= 0 && jv < nb && pos[jv] >= 0, FEEC_E_DEVICE, "coarsest stencil leaves its box");
      dense[r * nl + pos[jv]] += a;
    }
  }
  FEEC_REQUIRE(spd_inverse(nl, dense), FEEC_E_NOT_SPD, "coarsest multigrid operator is not SPD");
  std::vector<double> boxinv((size_t)nb * nb, 0.0);
  for (int64_t r = 0; r < nl; ++r)
    for (int64_t c = 0; c < nl; ++c) boxinv[(size_t)live[r] * nb + live[c]] = dense[r * nl + c];
  copy_to_device(ctx, h.cinv.p, boxinv.data(), boxinv.size() * 8);
  to32(ctx, h.cinv.p, (int64_t)boxinv.size(), h.cinv32);
  ctx->sync();
}

void alloc_slots(Ctx* ctx, Hx& h) {
  for (auto& l : h.L) {
    const int64_t m = std::max<int64_t>((int64_t)l.box.nb * h.nv, 1);
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
  if ((int64_t)h.s64.n != std::max<int64_t>(h.n, 1)) h.s64.alloc(ctx, std::max<int64_t>(h.n, 1));
  if (h.n) k_scale<<<grid_for(h.n), TPB, 0, ctx->stream>>>(s.inv_diag.p, h.n, kOmega, h.s64.p);
  FEEC_HIP(hipGetLastError());
  to32(ctx, h.s64.p, h.n, h.s32);
  for (auto& b : h.blocks) {
    if (b.j != 2) continue;
    // S_1 on the grade-1 DOFs: the sigma block's smoother
    const HxBlock* lower = nullptr;
    for (auto& o : h.blocks)
      if (o.j == 1) lower = &o;
    FEEC_REQUIRE(lower, FEEC_E_DEVICE, "grade-2 block without its grade-1 block");
    const int64_t nl = b.o_d.m.ncols;
    if ((int64_t)b.sl.n != std::max<int64_t>(nl, 1)) b.sl.alloc(ctx, std::max<int64_t>(nl, 1));
    FEEC_HIP(hipMemcpyAsync(b.sl.p, h.s64.p + lower->off, nl * 8, hipMemcpyDeviceToDevice, ctx->stream));
    to32(ctx, b.sl.p, nl, b.sl32);
  }
}

}  // namespace

namespace {
void hx_spectrum(Hx& h);

// The KKT preconditioner matrix P (blockdiag(hdif_gram(k-1), hdif_gram(k)))
// the Chebyshev steps multiply by, and its symmetric FP32 copy.
void bind_pc_matrix(Hx& h, const KktSystem& s) {
  h.P_rp = s.P.row_ptr.p;
  h.P_col = s.P.col.p;
  h.P_val = s.P.val.p;
  h.P_nnz = s.P.nnz;
  if (h.cheb_m > 1 && h.P_tpos.p && s.P.nnz)
    k_sym_single<<<grid_for(s.P.nnz), TPB, 0, h.ctx->stream>>>(s.P.val.p, h.P_tpos.p, s.P.nnz, h.P32.p);
  FEEC_HIP(hipGetLastError());
}

// The pattern side of P's FP32 copy: the transpose entry of every entry (P is
// structurally symmetric with sorted rows, so the counting-sort transpose
// lists (j, i) exactly where P's row j holds it) and, when every column
// lies within +-32767 of its row, the 16-bit column offsets.
void build_pc_single(Hx& h, const KktSystem& s) {
  Ctx* ctx = h.ctx;
  const int64_t n = s.P.nrows, nnz = s.P.nnz;
  h.P_tpos.alloc(ctx, std::max<int64_t>(nnz, 1));
  h.P32.alloc(ctx, std::max<int64_t>(nnz, 1));
  h.P_col16.reset();
  if (!nnz) return;
  DBuf<int16_t> c16(ctx, nnz);
  // flags[0]: some entry has no transpose; flags[1]: some offset is beyond 16 bits
  DBuf<int> flags(ctx, 2);
  k_zero_i32<<<1, 64, 0, ctx->stream>>>(flags.p, 2);
  k_pc_pattern<<<grid_for(nnz), TPB, 0, ctx->stream>>>(s.P.row_ptr.p, s.P.col.p, n, nnz, h.P_tpos.p, c16.p,
                                                        flags.p);
  FEEC_HIP(hipGetLastError());
  int hf[2] = {0, 0};
  copy_to_host(ctx, hf, flags.p, sizeof(hf));
  FEEC_REQUIRE(hf[0] == 0, FEEC_E_DEVICE, "preconditioner matrix is not structurally symmetric");
  if (hf[1] == 0) h.P_col16 = std::move(c16);
}
}  // namespace

Hx* hx_create(Geom& g, int bc, int k, const KktSystem& s, int sweeps, int cheb_m, double frac,
              bool spectrum_fp32, bool use_graph, std::string* why) {
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
  const int d = m.d;
  if (ipow(n + 1, d) * MAXNV >= ((int64_t)1 << 31)) return no("the grid is too large for 32-bit slot indices");
  auto h = std::make_unique<Hx>();
  h->ctx = ctx;
  h->d = d;
  h->k = k;
  h->bc = bc;
  h->grid = n;
  h->sweeps = std::max(1, sweeps);
  h->n = s.n;
  h->N = s.n + s.nh;
  h->split = s.split;
  h->cheb_m = std::max(1, cheb_m);
  h->frac = frac;
  h->spectrum_fp32 = spectrum_fp32;
  h->use_graph = use_graph;
  if (h->cheb_m > 1 && s.P.nnz == 0) return no("no preconditioner matrix for the Chebyshev steps");
  if (!(frac > 0.0 && frac < 1.0)) return no("the Chebyshev interval fraction must lie in (0, 1)");
  const bool rel = bc == FEEC_RELATIVE;
  if (rel) m.ensure_interior();
  const int NS = nstencil(d);
  // vertex -> its grade-0 DOF (-1: none)
  std::vector<int32_t> dof0(m.n[0]);
  if (rel)
    copy_to_host(ctx, dof0.data(), m.interior[0].p, m.n[0] * 4);
  else
    for (int64_t v = 0; v < m.n[0]; ++v) dof0[v] = (int32_t)v;
  // the levels
  {
    auto add_level = [&](int gcells, const std::vector<uint8_t>& al, int64_t nlive) {
      MgLevel l;
      l.box.g = gcells;
      l.box.nb = (int)ipow(gcells + 1, d);
      int st = 1;
      for (int a = 0; a < d; ++a) {
        l.box.st[a] = st;
        st *= gcells + 1;
      }
      l.live = nlive;
      l.alive.alloc(ctx, l.box.nb);
      copy_to_device(ctx, l.alive.p, al.data(), l.box.nb);
      l.A.alloc(ctx, (size_t)NS * l.box.nb);
      l.dinv.alloc(ctx, l.box.nb);
      l.bad.alloc(ctx, l.box.nb);
      h->L.push_back(std::move(l));
    };
    std::vector<uint8_t> alive(ipow(n + 1, d));
    int64_t live = 0;
    for (size_t v = 0; v < alive.size(); ++v) {
      alive[v] = dof0[v] >= 0;
      live += alive[v];
    }
    add_level(n, alive, live);
    int gr = n;
    while (live > kCoarseMax && gr >= 2) {
      const int gc = (gr + 1) / 2;
      std::vector<uint8_t> ac(ipow(gc + 1, d));
      live = 0;
      for (size_t w = 0; w < ac.size(); ++w) {
        bool in = true;
        int64_t r = (int64_t)w;
        for (int a = 0; a < d; ++a, r /= (gc + 1)) {
          const int64_t c = r % (gc + 1);
          if (rel) in = in && c > 0 && 2 * c <= gr - 1;
        }
        ac[w] = in;
        live += in;
      }
      add_level(gc, ac, live);
      gr = gc;
    }
    if (h->L.back().box.nb > kDenseMaxBox) return no("the coarsest multigrid level is too large");
  }
  // the level-0 operator's stencil positions, then the hierarchy's values
  {
    const Box& b0 = h->L[0].box;
    // the vertex of each nodal DOF (null: the DOF is the vertex)
    DBuf<int32_t> box_of;
    if (rel) {
      box_of.alloc(ctx, std::max<int64_t>(m.n_interior[0], 1));
      if (m.n[0])
        k_invert_map<<<grid_for(m.n[0]), TPB, 0, ctx->stream>>>(m.interior[0].p, m.n[0], m.n_interior[0],
                                                                 box_of.p);
      FEEC_HIP(hipGetLastError());
    }
    Csr A0;
    const Csr* src_op = &s.P;
    int64_t lo = 0, hi = 0;
    if (k <= 1 && s.P.nnz) {  // the nodal operator is P's block [lo, hi)
      lo = k == 1 ? 0 : s.split;
      hi = k == 1 ? s.split : s.n;
      h->l0_from_p = true;
    } else {
      nodal_operator(g, bc, s, k, A0);
      src_op = &A0;
      hi = A0.nrows;
    }
    const int64_t nn = hi - lo;
    FEEC_REQUIRE(nn == (rel ? m.n_interior[0] : m.n[0]), FEEC_E_DEVICE, "nodal block of the wrong size");
    DBuf<unsigned long long> cnt(ctx, nn + 1);
    DBuf<int64_t> start(ctx, nn + 1);
    k_l0_count<<<grid_for(nn + 1), TPB, 0, ctx->stream>>>(src_op->row_ptr.p, src_op->col.p, lo, hi, cnt.p);
    FEEC_HIP(hipGetLastError());
    exclusive_scan_counts(ctx, cnt.p, start.p, nn + 1);
    copy_to_host(ctx, &h->l0_nnz, start.p + nn, 8);
    h->l0_map.alloc(ctx, std::max<int64_t>(h->l0_nnz, 1));
    if (h->l0_from_p) h->l0_src.alloc(ctx, std::max<int64_t>(h->l0_nnz, 1));
    DBuf<int> bad(ctx, 1);
    k_zero_i32<<<1, 64, 0, ctx->stream>>>(bad.p, 1);
    if (nn) {
      const unsigned gl = grid_for(nn);
      int32_t* srcp = h->l0_from_p ? h->l0_src.p : nullptr;
#define FEEC_L0(D)                                                                                       \
  k_l0_fill<D><<<gl, TPB, 0, ctx->stream>>>(src_op->row_ptr.p, src_op->col.p, lo, hi, box_of.p, b0, start.p, \
                                            h->l0_map.p, srcp, bad.p)
      switch (d) {
        case 1: FEEC_L0(1); break;
        case 2: FEEC_L0(2); break;
        case 3: FEEC_L0(3); break;
        default: FEEC_L0(4); break;
      }
#undef FEEC_L0
      FEEC_HIP(hipGetLastError());
    }
    int hbad = 0;
    copy_to_host(ctx, &hbad, bad.p, sizeof(int));
    if (hbad) return no("the nodal operator is not a Kuhn stencil");
    const int nbc = h->L.back().box.nb;
    h->cinv.alloc(ctx, (size_t)nbc * nbc);
    try {
      hierarchy_values(*h, h->l0_from_p ? s.P.val.p : A0.val.p, true);
    } catch (const Error& e) {
      return no(e.what());
    }
  }
  // blocks: sigma (grade k-1, when k >= 1) and u (grade k)
  const int64_t ns = s.split;
  const int64_t nb0 = h->L[0].box.nb;
  std::vector<int> grades;
  if (k >= 1) grades.push_back(k - 1);
  grades.push_back(k);
  int slot = 0;
  // the vertices' box columns (null: every vertex keeps its own) and coordinates
  DBuf<int32_t> live_dev;
  const int32_t* live_col_dev = nullptr;
  if (rel) {
    live_dev.alloc(ctx, std::max<int64_t>(m.n[0], 1));
    k_live_col<<<grid_for(m.n[0]), TPB, 0, ctx->stream>>>(m.interior[0].p, m.n[0], live_dev.p);
    FEEC_HIP(hipGetLastError());
    live_col_dev = live_dev.p;
  }
  const double* X = k >= 1 ? device_coords(g) : nullptr;
  for (int j : grades) {
    HxBlock b;
    b.j = j;
    b.off = (j == k) ? ns : 0;
    b.n = (j == k) ? s.n - ns : ns;
    // the block's DOF of each j-simplex (null: identity) and its inverse
    const int32_t* map_j = rel ? m.interior[j].p : nullptr;
    if (j == 0) {
      HostCsr o;  // block DOF -> its vertex
      o.nrows = b.n;
      o.ncols = nb0;
      o.rp.resize(b.n + 1);
      o.col.assign(b.n, 0);
      o.val.assign(b.n, 1.0);
      for (int64_t v = 0; v < m.n[0]; ++v)
        if (dof0[v] >= 0 && dof0[v] < b.n) o.col[dof0[v]] = (int32_t)v;
      for (int64_t i = 0; i <= b.n; ++i) o.rp[i] = i;
      upload(ctx, transpose(o), b.g_main);
      upload(ctx, o, b.o_main);
      b.slot_main = slot;
      slot += 1;
    } else if (j == 1) {
      DBuf<int32_t> order;
      row_order(m, map_j, 1, b.n, order);
      XCsr D, Pi, DP;
      coboundary_dev(m, 1, map_j, b.n, live_col_dev, nb0, D);
      vector_nodal_dev(m, X, 1, order.p, live_col_dev, b.n, nb0, Pi);
      DBuf<int> bad(ctx, 2);
      k_zero_i32<<<1, 64, 0, ctx->stream>>>(bad.p, 2);
      interleave_dev(ctx, D, Pi, DP, bad.p);
      transpose_dev(ctx, DP, b.g_main);
      // the matrix-free scatter: ends and Pi values per row, after checking
      // that every D_0 entry is -1 at the smaller end and +1 at the larger
      b.ev.alloc(ctx, std::max<int64_t>(2 * b.n, 1));
      b.eh64.alloc(ctx, std::max<int64_t>(b.n * d, 1));
      if (b.n)
        k_edge_ends<<<grid_for(b.n), TPB, 0, ctx->stream>>>(m.faces[0].p, m.rep[1].p, d, order.p, b.n,
                                                            D.m.row_ptr.p, D.m.col.p, D.m.val.p, X, b.ev.p,
                                                            b.eh64.p, bad.p + 1);
      FEEC_HIP(hipGetLastError());
      int hb[2] = {0, 0};
      copy_to_host(ctx, hb, bad.p, sizeof(hb));
      FEEC_REQUIRE(hb[0] == 0, FEEC_E_DEVICE, "transfer patterns differ");
      b.slot_main = slot;
      slot += DP.C;
      if (hb[1] == 0 && d <= 4) {
        b.edge_free = true;
        to32(ctx, b.eh64.p, b.n * d, b.eh32);
      } else {
        b.ev.reset();
        b.eh64.reset();
        b.o_main = std::move(DP);
      }
    } else {
      const int32_t* map_l = rel ? m.interior[1].p : nullptr;
      const int64_t nl = rel ? m.n_interior[1] : m.n[1];
      DBuf<int32_t> order_j, order_l;
      row_order(m, map_j, 2, b.n, order_j);
      row_order(m, map_l, 1, nl, order_l);
      coboundary_dev(m, 2, map_j, b.n, map_l, nl, b.o_d);
      vector_nodal_dev(m, X, 2, order_j.p, live_col_dev, b.n, nb0, b.o_main);
      vector_nodal_dev(m, X, 1, order_l.p, live_col_dev, nl, nb0, b.o_ul);
      transpose_dev(ctx, b.o_d, b.g_t);
      transpose_dev(ctx, b.o_ul, b.g_tl);
      transpose_dev(ctx, b.o_main, b.g_main);
      b.slot_main = slot;
      slot += b.o_main.C;
      b.slot_tl = slot;
      slot += b.o_ul.C;
      b.t64.alloc(ctx, std::max<int64_t>(nl, 1));
      b.u64.alloc(ctx, std::max<int64_t>(nl, 1));
      b.t32.alloc(ctx, std::max<int64_t>(nl, 1));
      b.u32.alloc(ctx, std::max<int64_t>(nl, 1));
    }
    h->blocks.push_back(std::move(b));
  }
  if (slot > MAXNV) return no("too many auxiliary copies for one cycle");
  for (const auto& b : h->blocks)
    for (const XCsr* x : {&b.g_main, &b.o_main, &b.g_t, &b.g_tl, &b.o_ul, &b.o_d})
      if (x->C > MAXC) return no("too many components per transfer entry");
  h->nv = slot;
  alloc_slots(ctx, *h);
  h->zt64.alloc(ctx, std::max<int64_t>(h->N, 1));
  h->zt32.alloc(ctx, std::max<int64_t>(h->N, 1));
  smoother_diag(*h, s);
  // the Chebyshev steps
  if (h->cheb_m > 1 && s.P.nnz) {
    try {
      build_pc_single(*h, s);
    } catch (const Error& e) {
      return no(e.what());
    }
  }
  bind_pc_matrix(*h, s);
  if (h->cheb_m > 1) {
    const int64_t nn = std::max<int64_t>(h->n, 1);
    for (DBuf<double>* b : {&h->res64, &h->d64, &h->zc64, &h->w64, &h->pvec}) b->alloc(ctx, nn);
    for (DBuf<float>* b : {&h->res32, &h->d32, &h->zc32, &h->w32}) b->alloc(ctx, nn);
    h->coef.alloc(ctx, 2 + 3 * (h->cheb_m - 1));
    h->part.alloc(ctx, 4 * (size_t)NBLK);
    h->lz.alloc(ctx, 14 + 4 * kMaxLanczos);
    try {
      hx_spectrum(*h);
    } catch (const Error& e) {
      return no(e.what());
    }
  }
  // work estimate of one application in FP32
  int64_t bytes = 0, launches = 1;
  for (size_t i = 0; i < h->L.size(); ++i) {
    const MgLevel& l = h->L[i];
    const int64_t nbv = (int64_t)l.box.nb * h->nv;
    if (i + 1 < h->L.size()) {
      bytes += (int64_t)(2 * h->sweeps) * ((int64_t)(NS / 2 + 1) * l.box.nb * 4 + 12 * nbv) + 16 * nbv;
      launches += 2 + 2 * h->sweeps;
    } else {
      bytes += (int64_t)l.box.nb * l.box.nb * 4;
      launches += 1;
    }
  }
  for (const auto& b : h->blocks)
    for (const XCsr* x : {&b.g_main, &b.o_main, &b.g_t, &b.g_tl, &b.o_ul, &b.o_d})
      if (x->m.nnz) {
        bytes += x->m.nnz * 4 * (1 + x->C) + x->m.nrows * 8;
        launches += 1;
      }
  if (h->cheb_m > 1) {
    bytes = bytes * h->cheb_m + (h->cheb_m - 1) * (s.P.nnz * (h->P_col16.p ? 6 : 8) + h->n * 28) + h->n * 12;
    launches = launches * h->cheb_m + 2 * (h->cheb_m - 1) + 2;
  }
  h->bytes32 = bytes;
  h->launches = launches;
  ctx->sync();
  return h.release();
}


void hx_refresh(Hx& h, Geom& g, const KktSystem& s) {
  FEEC_REQUIRE(s.P.nnz == h.P_nnz, FEEC_E_DEVICE, "preconditioner matrix pattern changed");
  if (h.l0_from_p) {
    hierarchy_values(h, s.P.val.p, false);
  } else {
    Csr A0;
    nodal_operator(g, h.bc, s, h.k, A0);
    FEEC_REQUIRE(A0.nnz == h.l0_nnz, FEEC_E_DEVICE, "nodal operator pattern changed");
    hierarchy_values(h, A0.val.p, false);
  }
  smoother_diag(h, s);
  bind_pc_matrix(h, s);
  hx_spectrum(h);
  h.ctx->sync();
}

// Largest eigenvalue of the symmetric tridiagonal (diag, off): bisection on
// Sturm counts inside the Gershgorin bound.
double tridiagonal_lmax(const std::vector<double>& diag, const std::vector<double>& off) {
  const int n = (int)diag.size();
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < n; ++i) {
    const double r = (i ? std::fabs(off[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(off[i]) : 0.0);
    lo = std::min(lo, diag[i] - r);
    hi = std::max(hi, diag[i] + r);
  }
  for (int it = 0; it < 60; ++it) {
    const double mid = 0.5 * (lo + hi);
    int count = 0;
    double q = 1.0;
    for (int i = 0; i < n; ++i) {  // eigenvalues below mid = negative pivots of T - mid I
      q = diag[i] - mid - (i ? off[i - 1] * off[i - 1] / q : 0.0);
      if (q == 0.0) q = -1e-300;
      count += q < 0.0;
    }
    if (count == n)
      hi = mid;
    else
      lo = mid;
  }
  return hi;
}

namespace {

template <class TP>
struct Sel;
template <>
struct Sel<double> {
  static const double* A(const MgLevel& l) { return l.A.p; }
  static const double* dinv(const MgLevel& l) { return l.dinv.p; }
  static double* r(MgLevel& l) { return l.r64.p; }
  static double* x(MgLevel& l) { return l.x64.p; }
  static double* y(MgLevel& l) { return l.y64.p; }
  static double* e(MgLevel& l) { return l.e64.p; }
  static const double* cinv(const Hx& h) { return h.cinv.p; }
  static const double* s(const Hx& h) { return h.s64.p; }
  static double* zt(Hx& h) { return h.zt64.p; }
  static const double* v(const XCsr& x) { return x.m.val.p; }
  static const double* sl(const HxBlock& b) { return b.sl.p; }
  static double* t(HxBlock& b) { return b.t64.p; }
  static double* u(HxBlock& b) { return b.u64.p; }
  static const double* eh(const HxBlock& b) { return b.eh64.p; }
  static double* res(Hx& h) { return h.res64.p; }
  static double* d(Hx& h) { return h.d64.p; }
  static double* zc(Hx& h) { return h.zc64.p; }
  static double* w(Hx& h) { return h.w64.p; }
};
template <>
struct Sel<float> {
  static const float* A(const MgLevel& l) { return l.A32.p; }
  static const float* dinv(const MgLevel& l) { return l.dinv32.p; }
  static float* r(MgLevel& l) { return l.r32.p; }
  static float* x(MgLevel& l) { return l.x32.p; }
  static float* y(MgLevel& l) { return l.y32.p; }
  static float* e(MgLevel& l) { return l.e32.p; }
  static const float* cinv(const Hx& h) { return h.cinv32.p; }
  static const float* s(const Hx& h) { return h.s32.p; }
  static float* zt(Hx& h) { return h.zt32.p; }
  static const float* v(const XCsr& x) { return x.v32.p; }
  static const float* sl(const HxBlock& b) { return b.sl32.p; }
  static float* t(HxBlock& b) { return b.t32.p; }
  static float* u(HxBlock& b) { return b.u32.p; }
  static const float* eh(const HxBlock& b) { return b.eh32.p; }
  static float* res(Hx& h) { return h.res32.p; }
  static float* d(Hx& h) { return h.d32.p; }
  static float* zc(Hx& h) { return h.zc32.p; }
  static float* w(Hx& h) { return h.w32.p; }
};

// One V-cycle on level i with its rhs in r(i); returns the buffer holding
// the result (x(i) or y(i)).
template <class TP, int D, bool FINE>
void sweep_launch(Enq& q, int mode, const MgLevel& l, const TP* A, const TP* dinv, const TP* r, const TP* x,
                  TP* out, unsigned nv, const MinresS* S) {
  const unsigned g = grid_for((int64_t)l.box.nb * nv);
  switch (mode) {
    case 0: q(k_st_sweep<TP, D, 0, FINE>, g, TPB, A, dinv, r, x, out, l.box, nv, S); break;
    case 1: q(k_st_sweep<TP, D, 1, FINE>, g, TPB, A, dinv, r, x, out, l.box, nv, S); break;
    default: q(k_st_sweep<TP, D, 2, FINE>, g, TPB, A, dinv, r, x, out, l.box, nv, S); break;
  }
}

// One V-cycle on level i with its rhs in r(i); returns the buffer holding
// the result (x(i) or y(i)).
template <class TP, int D>
TP* vcycle(Enq& q, Hx& h, size_t i, const MinresS* S) {
  using X = Sel<TP>;
  MgLevel& l = h.L[i];
  const unsigned nv = (unsigned)h.nv;
  const int64_t nbv = (int64_t)l.box.nb * nv;
  if (i + 1 == h.L.size()) {
    q(k_st_dense<TP>, (unsigned)l.box.nb, TPB, X::cinv(h), (const TP*)X::r(l), X::x(l), l.box.nb, h.nv, S);
    return X::x(l);
  }
  MgLevel& c = h.L[i + 1];
  TP* x = X::x(l);
  TP* y = X::y(l);
  const TP* r = X::r(l);
  const TP* A = X::A(l);
  const TP* dinv = X::dinv(l);
  auto sweep = [&](int mode, TP* out) {
    if (i == 0)
      sweep_launch<TP, D, true>(q, mode, l, A, dinv, r, x, out, nv, S);
    else
      sweep_launch<TP, D, false>(q, mode, l, A, dinv, r, x, out, nv, S);
  };
  // pre-smoothing from x = 0: x = dinv r, then sweeps - 1 sweeps (the first fused with it)
  if (h.sweeps == 1) {
    q(k_st_scale<TP>, grid_for(nbv), TPB, dinv, r, x, (unsigned)l.box.nb, nv, S);
  } else {
    sweep(2, y);
    std::swap(x, y);
    for (int s = 2; s < h.sweeps; ++s) {
      sweep(0, y);
      std::swap(x, y);
    }
  }
  sweep(1, X::e(l));
  q(k_st_restrict<TP, D>, grid_for((int64_t)c.box.nb * nv), TPB, (const TP*)X::e(l), X::r(c),
    (const uint8_t*)c.alive.p, l.box, c.box, nv, S);
  TP* xc = vcycle<TP, D>(q, h, i + 1, S);
  q(k_st_prolong<TP, D>, grid_for(nbv), TPB, (const TP*)xc, x, dinv, l.box, c.box, nv, S);
  for (int s = 0; s < h.sweeps; ++s) {
    sweep(0, y);
    std::swap(x, y);
  }
  return x;
}

template <class TX, class TP>
void gather(Enq& q, const XCsr& g, const TX* x, TP* y, int nv, int slot0, const MinresS* S) {
  using X = Sel<TP>;
  const int64_t nr = g.m.nrows;
  if (g.avg_row >= 6.0)
    q(k_hx_gather<TX, TP, 4>, grid_for(nr * 4), TPB, g.m.row_ptr.p, g.m.col.p, X::v(g), g.C, x, y, nv, slot0,
      nr, S);
  else
    q(k_hx_gather<TX, TP, 1>, grid_for(nr), TPB, g.m.row_ptr.p, g.m.col.p, X::v(g), g.C, x, y, nv, slot0, nr,
      S);
}

// zt = B r on the preconditioned rows (r of length n at least).
template <class TIN, class TP, int D>
void hx_core_d(Enq& q, Hx& h, const TIN* r, const MinresS* S) {
  using X = Sel<TP>;
  MgLevel& l0 = h.L[0];
  TP* slots = X::r(l0);
  TP* zt = X::zt(h);
  const int nv = h.nv;
  for (HxBlock& b : h.blocks) {
    const TIN* rb = r + b.off;
    if (b.j == 2) {
      gather<TIN, TP>(q, b.g_t, rb, X::t(b), 1, 0, S);
      gather<TP, TP>(q, b.g_tl, (const TP*)X::t(b), slots, nv, b.slot_tl, S);
    }
    gather<TIN, TP>(q, b.g_main, rb, slots, nv, b.slot_main, S);
  }
  const TP* xs = vcycle<TP, D>(q, h, 0, S);
  for (HxBlock& b : h.blocks) {
    TP* zb = zt + b.off;
    const TIN* rb = r + b.off;
    const TP* sb = X::s(h) + b.off;
    const XCsr& om = b.o_main;
    if (b.j == 2) {
      const XCsr& ul = b.o_ul;
      q(k_hx_out<TP, TP, TP>, grid_for(ul.m.nrows), TPB, ul.m.row_ptr.p, ul.m.col.p, X::v(ul), ul.C, 0, xs,
        (const TP*)slots, X::dinv(l0), nv, b.slot_tl, X::sl(b), (const TP*)X::t(b), X::u(b), ul.m.nrows, 0, S);
      const XCsr& od = b.o_d;
      q(k_hx_out<TP, TP, TIN>, grid_for(od.m.nrows), TPB, od.m.row_ptr.p, od.m.col.p, X::v(od), od.C, 0,
        (const TP*)X::u(b), (const TP*)nullptr, (const TP*)nullptr, 1, 0, sb, rb, zb, od.m.nrows, 0, S);
      q(k_hx_out<TP, TP, TIN>, grid_for(om.m.nrows), TPB, om.m.row_ptr.p, om.m.col.p, X::v(om), om.C, 0, xs,
        (const TP*)slots, X::dinv(l0), nv, b.slot_main, (const TP*)nullptr, (const TIN*)nullptr, zb,
        om.m.nrows, 1, S);
    } else if (b.edge_free) {
      q(k_hx_out_edge<TP, TIN, D>, grid_for(b.n), TPB, (const int32_t*)b.ev.p, X::eh(b), xs, (const TP*)slots,
        X::dinv(l0), nv, b.slot_main, sb, rb, zb, b.n, S);
    } else {
      q(k_hx_out<TP, TP, TIN>, grid_for(om.m.nrows), TPB, om.m.row_ptr.p, om.m.col.p, X::v(om), om.C,
        b.j == 1 ? 1 : 0, xs, (const TP*)slots, X::dinv(l0), nv, b.slot_main, sb, rb, zb, om.m.nrows, 0, S);
    }
  }
}

// y = P x: FP64 values on the FP64 path, the symmetric FP32 copy otherwise.
template <class TP>
void pc_spmv(Enq& q, Hx& h, const TP* x, TP* y, const MinresS* S) {
  const unsigned g = grid_for(h.n * 4);
  if constexpr (std::is_same_v<TP, double>) {
    q(k_pc_spmv<double, double, 4, 2>, g, TPB, h.P_rp, h.P_col, h.P_val, x, y, h.n, S);
  } else {
    if (h.P_col16.p)
      q(k_pc_spmv32<int16_t, 4, 2>, g, TPB, h.P_rp, (const int16_t*)h.P_col16.p, (const float*)h.P32.p, x, y,
        h.n, S);
    else
      q(k_pc_spmv32<int32_t, 4, 2>, g, TPB, h.P_rp, h.P_col, (const float*)h.P32.p, x, y, h.n, S);
  }
}

// z = B r (cheb_m = 1) or p(B P) B r per block, z = r on the border rows,
// and the partial sums of r . z.
template <class TIN, class TP, class TOUT, int D>
void hx_apply_d(Enq& q, Hx& h, const TIN* r, TOUT* z, const MinresS* S, double* part) {
  using X = Sel<TP>;
  hx_core_d<TIN, TP, D>(q, h, r, S);
  if (h.cheb_m <= 1) {
    q(k_hx_final<TIN, TP, TOUT>, red_grid(h.N), TPB, r, (const TP*)X::zt(h), z, h.n, h.N, S, part);
    return;
  }
  TP* res = X::res(h);
  TP* d = X::d(h);
  TP* zc = X::zc(h);
  TP* w = X::w(h);
  const double* coef = h.coef.p;
  q(k_hxc_first<TP>, grid_for(h.n), TPB, (const TP*)X::zt(h), res, d, zc, h.n, h.split, coef, S);
  for (int k = 0; k + 1 < h.cheb_m; ++k) {
    pc_spmv<TP>(q, h, (const TP*)d, w, S);
    hx_core_d<TP, TP, D>(q, h, (const TP*)w, S);
    q(k_hxc_step<TP>, grid_for(h.n), TPB, (const TP*)X::zt(h), res, d, zc, h.n, h.split, coef, k, S);
  }
  q(k_hxc_final<TIN, TP, TOUT>, red_grid(h.N), TPB, r, (const TP*)zc, (const TP*)d, z, h.n, h.N, S, part);
}

template <class TIN, class TP, class TOUT>
void hx_apply(Enq& q, Hx& h, const TIN* r, TOUT* z, const MinresS* S, double* part) {
  switch (h.d) {
    case 1: hx_apply_d<TIN, TP, TOUT, 1>(q, h, r, z, S, part); break;
    case 2: hx_apply_d<TIN, TP, TOUT, 2>(q, h, r, z, S, part); break;
    case 3: hx_apply_d<TIN, TP, TOUT, 3>(q, h, r, z, S, part); break;
    default: hx_apply_d<TIN, TP, TOUT, 4>(q, h, r, z, S, part); break;
  }
}

void hx_core(Enq& q, Hx& h, const double* r) {
  switch (h.d) {
    case 1: hx_core_d<double, double, 1>(q, h, r, nullptr); break;
    case 2: hx_core_d<double, double, 2>(q, h, r, nullptr); break;
    case 3: hx_core_d<double, double, 3>(q, h, r, nullptr); break;
    default: hx_core_d<double, double, 4>(q, h, r, nullptr); break;
  }
}

// lambda_max of B P per block: kLanczosSteps steps of CG on P preconditioned
// by B (both blocks at once, each with its own scalars) from the same
// deterministic start every time, and the largest eigenvalue of the Lanczos
// tridiagonal their coefficients make; then the Chebyshev coefficients.
constexpr int kLanczosSteps = 6;
// a block's CG stops early once r.Br falls below this fraction of its start,
// near the roundoff floor of the arithmetic (the estimate from fewer steps is
// still a Ritz value, a lower bound)
template <class TP>
constexpr double kLanczosStop = std::is_same_v<TP, float> ? 1e-10 : 1e-20;
constexpr double kLmaxMargin = 1.05;


// kLanczosSteps of B-preconditioned CG on P per block in precision TP; the
// largest eigenvalue of each block's Lanczos tridiagonal into h.lam.
// The estimate's launches: start vector, B r, then kLanczosSteps steps; every
// scalar stays on the device (lz), so the whole sequence is one graph.
template <class TP, int D>
void lanczos_enqueue(Enq& q, Hx& h) {
  using X = Sel<TP>;
  const int64_t n = h.n, split = h.split;
  TP* r = X::res(h);
  TP* pd = X::d(h);
  TP* w = X::w(h);
  TP* z = X::zt(h);
  const int64_t lo[2] = {0, split}, hi[2] = {split, n};
  const int g[2] = {split > 0 ? (int)red_grid(hi[0] - lo[0]) : 0, n > split ? (int)red_grid(hi[1] - lo[1]) : 0};
  auto dots = [&](const TP* x, const TP* y) {
    for (int b = 0; b < 2; ++b)
      if (g[b]) q(k_dotr<TP>, (unsigned)g[b], TPB, x, y, lo[b], hi[b], h.part.p + b * NBLK);
  };
  q(k_convert<double, TP>, grid_for(n), TPB, (const double*)h.pvec.p, r, n);
  hx_core_d<TP, TP, D>(q, h, (const TP*)r, nullptr);
  q(k_convert<TP, TP>, grid_for(n), TPB, (const TP*)z, pd, n);
  dots(r, z);
  q(k_lz_init, 1, 64, (const double*)h.part.p, g[0], g[1], h.lz.p);
  for (int j = 0; j < kLanczosSteps; ++j) {
    pc_spmv<TP>(q, h, (const TP*)pd, w, nullptr);
    dots(pd, w);
    q(k_lz_alpha, 1, 64, (const double*)h.part.p, g[0], g[1], h.lz.p);
    q(k_lz_axpy<TP>, grid_for(n), TPB, (const TP*)w, r, n, split, (const double*)h.lz.p);
    hx_core_d<TP, TP, D>(q, h, (const TP*)r, nullptr);
    dots(r, z);
    q(k_lz_beta, 1, 64, (const double*)h.part.p, g[0], g[1], kLanczosStop<TP>, h.lz.p);
    q(k_lz_xpby<TP>, grid_for(n), TPB, (const TP*)z, pd, n, split, (const double*)h.lz.p);
  }
}

template <class TP, int D>
void lanczos(Hx& h) {
  static_assert(kLanczosSteps <= kMaxLanczos, "Lanczos history size");
  Ctx* ctx = h.ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = h.n, split = h.split;
  if (!h.lz_exec && !h.lz_direct_ready) {
    // the start vector, 0.5 + U[0,1) from xorshift64 (13, 7, 17), the same every time
    xorshift_start(ctx, h.pvec.p, n);
    if (h.use_graph) {
      FEEC_HIP(hipGraphCreate(&h.lz_graph, 0));
      Enq q{st, h.lz_graph, &h.lz_keep, nullptr};
      lanczos_enqueue<TP, D>(q, h);
      FEEC_HIP(hipGraphInstantiate(&h.lz_exec, h.lz_graph, nullptr, nullptr, 0));
    }
    h.lz_direct_ready = true;
  }
  if (h.lz_exec) {
    FEEC_HIP(hipGraphLaunch(h.lz_exec, st));
  } else {  // the same launches one by one (for tracing tools)
    Enq q{st};
    lanczos_enqueue<TP, D>(q, h);
  }
  std::vector<double> lz(14 + 4 * kMaxLanczos);
  copy_to_host(ctx, lz.data(), h.lz.p, lz.size() * 8);
  const bool has[2] = {split > 0, n > split};
  for (int b = 0; b < 2; ++b) {
    h.lam[b] = 0.0;
    if (!has[b]) continue;
    const int m = (int)lz[10 + b];
    FEEC_REQUIRE(m > 0, FEEC_E_NOT_SPD, "multigrid preconditioner spectrum estimate failed");
    auto al = [&](int j) { return lz[14 + 2 * j + b]; };
    auto be = [&](int j) { return lz[14 + 2 * kMaxLanczos + 2 * j + b]; };
    std::vector<double> diag(m), off(std::max(m - 1, 0));
    for (int j = 0; j < m; ++j) diag[j] = 1.0 / al(j) + (j ? be(j - 1) / al(j - 1) : 0.0);
    for (int j = 0; j + 1 < m; ++j) off[j] = std::sqrt(be(j)) / al(j);
    h.lam[b] = tridiagonal_lmax(diag, off);
  }
}

template <class TP>
void lanczos_d(Hx& h) {
  switch (h.d) {
    case 1: lanczos<TP, 1>(h); break;
    case 2: lanczos<TP, 2>(h); break;
    case 3: lanczos<TP, 3>(h); break;
    default: lanczos<TP, 4>(h); break;
  }
}

void hx_spectrum(Hx& h) {
  if (h.cheb_m <= 1 || h.n == 0) return;
  if (h.spectrum_fp32)
    lanczos_d<float>(h);
  else
    lanczos_d<double>(h);
  Ctx* ctx = h.ctx;
  // coefficients on [frac b, b], b = 1.05 lambda_max, per block
  std::vector<double> c(2 + 3 * (size_t)(h.cheb_m - 1));
  double theta[2], delta[2];
  for (int b = 0; b < 2; ++b) {
    const double bb = kLmaxMargin * (h.lam[b] > 0.0 ? h.lam[b] : 1.0), aa = h.frac * bb;
    theta[b] = 0.5 * (bb + aa);
    delta[b] = 0.5 * (bb - aa);
    c[b] = 1.0 / theta[b];
  }
  const double sigma1 = (1.0 + h.frac) / (1.0 - h.frac);  // theta / delta, both blocks
  double rk = 1.0 / sigma1;
  for (int k = 0; k + 1 < h.cheb_m; ++k) {
    const double rnext = 1.0 / (2.0 * sigma1 - rk);
    c[2 + 3 * k] = rnext * rk;
    c[3 + 3 * k] = 2.0 * rnext / delta[0];
    c[4 + 3 * k] = 2.0 * rnext / delta[1];
    rk = rnext;
  }
  copy_to_device(ctx, h.coef.p, c.data(), c.size() * 8);
  ctx->sync();
}

}  // namespace

// Mean seconds per launch, replayed from a graph of `reps` launches on the
// context's stream (or launched one by one): [0] a finest-level Jacobi sweep (FP32), [1] one block
// preconditioner application as an FP64_PC32 solve applies it (FP64 in and
// out, the cycles and Chebyshev steps in FP32), [2] the Chebyshev steps'
// product by the preconditioner matrix (FP32), [3] one B application (FP32).
void hx_kernel_times(Hx& h, int reps, bool use_graph, double out[4]) {
  Ctx* ctx = h.ctx;
  hipStream_t st = ctx->stream;
  reps = std::max(1, reps);
  MgLevel& l0 = h.L[0];
  const unsigned nv = (unsigned)h.nv;
  for (DBuf<float>* b : {&l0.r32, &l0.x32, &l0.y32})
    FEEC_HIP(hipMemsetAsync(b->p, 0, b->n * sizeof(float), st));
  DBuf<double> r(ctx, std::max<int64_t>(h.N, 1)), z(ctx, std::max<int64_t>(h.N, 1));
  DBuf<double> part(ctx, NBLK);
  DBuf<float> r32(ctx, std::max<int64_t>(h.N, 1));
  FEEC_HIP(hipMemsetAsync(r.p, 0, r.n * 8, st));
  FEEC_HIP(hipMemsetAsync(r32.p, 0, r32.n * 4, st));
  if (h.cheb_m > 1)
    for (DBuf<float>* b : {&h.d32, &h.w32}) FEEC_HIP(hipMemsetAsync(b->p, 0, b->n * 4, st));
  auto timed = [&](auto&& body) -> double {
    if (!use_graph) {  // launched one by one (for tracing and counter tools)
      Enq q{st};
      body(q);
      hipEvent_t e0, e1;
      FEEC_HIP(hipEventCreate(&e0));
      FEEC_HIP(hipEventCreate(&e1));
      FEEC_HIP(hipEventRecord(e0, st));
      for (int i = 0; i < reps; ++i) body(q);
      FEEC_HIP(hipEventRecord(e1, st));
      FEEC_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      FEEC_HIP(hipEventElapsedTime(&ms, e0, e1));
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return ms * 1e-3 / reps;
    }
    std::vector<std::shared_ptr<void>> keep;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    FEEC_HIP(hipGraphCreate(&graph, 0));
    double sec = 0.0;
    try {
      Enq q{st, graph, &keep, nullptr};
      for (int i = 0; i < reps; ++i) body(q);
      FEEC_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      FEEC_HIP(hipGraphLaunch(exec, st));  // warm
      hipEvent_t e0, e1;
      FEEC_HIP(hipEventCreate(&e0));
      FEEC_HIP(hipEventCreate(&e1));
      FEEC_HIP(hipEventRecord(e0, st));
      FEEC_HIP(hipGraphLaunch(exec, st));
      FEEC_HIP(hipEventRecord(e1, st));
      FEEC_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      FEEC_HIP(hipEventElapsedTime(&ms, e0, e1));
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      sec = ms * 1e-3 / reps;
    } catch (...) {
      if (exec) (void)hipGraphExecDestroy(exec);
      (void)hipGraphDestroy(graph);
      throw;
    }
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    return sec;
  };
  out[0] = h.L.size() > 1 ? timed([&](Enq& q) {
    switch (h.d) {
      case 1: sweep_launch<float, 1, true>(q, 0, l0, l0.A32.p, l0.dinv32.p, l0.r32.p, l0.x32.p, l0.y32.p, nv, nullptr); break;
      case 2: sweep_launch<float, 2, true>(q, 0, l0, l0.A32.p, l0.dinv32.p, l0.r32.p, l0.x32.p, l0.y32.p, nv, nullptr); break;
      case 3: sweep_launch<float, 3, true>(q, 0, l0, l0.A32.p, l0.dinv32.p, l0.r32.p, l0.x32.p, l0.y32.p, nv, nullptr); break;
      default: sweep_launch<float, 4, true>(q, 0, l0, l0.A32.p, l0.dinv32.p, l0.r32.p, l0.x32.p, l0.y32.p, nv, nullptr); break;
    }
  }) : 0.0;
  out[1] = timed([&](Enq& q) { hx_apply<double, float, double>(q, h, (const double*)r.p, z.p, nullptr, part.p); });
  const auto core = [&](Enq& q) {
    switch (h.d) {
      case 1: hx_core_d<float, float, 1>(q, h, (const float*)r32.p, nullptr); break;
      case 2: hx_core_d<float, float, 2>(q, h, (const float*)r32.p, nullptr); break;
      case 3: hx_core_d<float, float, 3>(q, h, (const float*)r32.p, nullptr); break;
      default: hx_core_d<float, float, 4>(q, h, (const float*)r32.p, nullptr); break;
    }
  };
  out[3] = timed(core);
  // the product by P as a Chebyshev step meets it: each launch timed on its own
  // events right after an (untimed) HX core has streamed its operands through
  // the caches. Back to back with itself the product finds part of P still
  // cached (about 20% faster at a 206 MB P), which a solve never sees.
  out[2] = 0.0;
  if (h.cheb_m > 1) {
    Enq q{st};
    hipEvent_t e0, e1;
    FEEC_HIP(hipEventCreate(&e0));
    FEEC_HIP(hipEventCreate(&e1));
    double sec = 0.0;
    for (int i = 0; i < reps; ++i) {
      core(q);
      FEEC_HIP(hipEventRecord(e0, st));
      pc_spmv<float>(q, h, (const float*)h.d32.p, h.w32.p, nullptr);
      FEEC_HIP(hipEventRecord(e1, st));
      FEEC_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      FEEC_HIP(hipEventElapsedTime(&ms, e0, e1));
      sec += ms * 1e-3;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    out[2] = sec / reps;
  }
}

void hx_info(const Hx& h, int* cheb_m, double lam[2], int* column_bytes) {
  if (cheb_m) *cheb_m = h.cheb_m;
  if (column_bytes) *column_bytes = h.P_col16.p ? 2 : 4;
  if (lam) {
    lam[0] = h.lam[0];
    lam[1] = h.lam[1];
  }
}

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
