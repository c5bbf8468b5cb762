// Krylov solvers on the device: CSR SpMV (FP64/FP32), preconditioned MINRES
// with every scalar of the recurrence kept in device memory, FP64 iterative
// refinement around an FP32 MINRES, and Jacobi-preconditioned CG.
//
// Reference behaviour (luiswirth/formoniq, crates/):
//   minres          iterative/src/krylov.rs:113-211 (Paige-Saunders, x0 = 0,
//                   stops on phibar/beta1 <= rtol)
//   cg              iterative/src/krylov.rs:48-95
//   Jacobi          iterative/src/precond.rs:53-123
//   BlockDiagonal   iterative/src/precond.rs:138-167
//   solve_source    formoniq/src/problems/elliptic.rs:132-182 (rtol 1e-10)
//
// One MINRES iteration is five launches that never return to the host:
//   spmv      yn = A (y/beta) - (beta/oldb) r1       + partial alpha = v.yn
//   alpha     alpha = sum of partials (fixed order)
//   update    yn -= (alpha/beta) r2, z = P yn         + partial beta^2 = yn.z
//   rotate    Givens rotation and w-update coefficients (one thread)
//   wx        w_new = (v - oldeps w2 - delta w)/gamma, x += phi w_new
// Reductions go through a fixed grid of per-block partials summed in a fixed
// order, so a run is bit-reproducible. The buffer roles rotate with period two
// (r1/r2, y/z, w/w2), so an even number of iterations is captured once into a
// hipGraph and replayed; a `done` flag in the scalar block turns the tail of a
// segment into no-ops once the tolerance is met, and the host reads the flag
// only between segments. The graph is built node by node (no stream capture),
// so solves on other threads' streams are never constrained by a capture.
//
// Mixed precision: the matrix values and every Krylov vector are FP32, SpMV
// accumulates in FP32; the orthogonalisation scalars (alpha, beta, the
// rotation) and their dot products are accumulated in FP64. The outer loop
// computes the residual b - A x with the FP64 matrix and FP64 x, and stops on
// the true relative residual.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <tuple>
#include <type_traits>
#include <vector>

#include "feec_internal.hpp"
#include "enqueue.hpp"

namespace feec {

namespace {

constexpr int TPB = 256;
constexpr int NBLK = 4096;  // fixed reduction grid: 16 blocks per CU on 256 CUs
constexpr int LANES = 8;    // lanes per CSR row in the KKT SpMV (about 28 nonzeros a row)
constexpr int PLANES = 4;   // lanes per row in the preconditioner SpMV (about 16 a row)

template <class T>
__device__ __forceinline__ double to_d(T v) {
  return (double)v;
}

// Dot product of CSR row `row` with x by L lanes of a wavefront, returned
// on every lane. Each lane keeps SPMV_UNROLL independent loads in flight; the
// summation order is fixed by the row pattern alone.
constexpr int SPMV_UNROLL = 2;
// Column of entry p of row `row`: stored directly (int32) or as a signed
// 16-bit offset from the row (a square matrix whose band fits in +-32767).
__device__ __forceinline__ int64_t col_at(const int32_t* __restrict__ col, int64_t p, int64_t) {
  return col[p];
}
__device__ __forceinline__ int64_t col_at(const int16_t* __restrict__ col, int64_t p, int64_t row) {
  return row + col[p];
}

template <class T, class I = int32_t, class V = T, int L = LANES>
__device__ __forceinline__ T row_dot(const int64_t* __restrict__ rp, const I* __restrict__ col,
                                     const V* __restrict__ val, const T* __restrict__ x, int64_t row,
                                     int lane) {
  constexpr int U = SPMV_UNROLL;
  const int64_t end = rp[row + 1];
  T a[U];
#pragma unroll
  for (int u = 0; u < U; ++u) a[u] = 0;
  int64_t p = rp[row] + lane;
  for (; p + (U - 1) * L < end; p += U * L) {
    int64_t c[U];
    V v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      c[u] = col_at(col, p + u * L, row);
      v[u] = val[p + u * L];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] += (T)v[u] * x[c[u]];
  }
#pragma unroll
  for (int u = 0; u + 1 < U; ++u)
    if (p + u * L < end) a[u] += (T)val[p + u * L] * x[col_at(col, p + u * L, row)];
  T acc = a[0];
#pragma unroll
  for (int u = 1; u < U; ++u) acc += a[u];
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, L);
  return acc;
}

__device__ __forceinline__ double block_reduce(double v, double* sh) {
  // wavefront (64-lane) tree, then across the 4 wavefronts
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += sh[w];
  }
  return r;  // valid on thread 0
}

// Workgroups are dispatched to the 8 XCDs of the chip round-robin (workgroup b
// to XCD b % 8), each XCD with its own L2. xcd_block() maps them so that XCD x
// runs the x-th contiguous eighth of the logical blocks: its L2 then fetches
// only the part of the gathered vector its rows touch, not all of it. The grid
// is a multiple of 8; logical blocks >= nblocks exit.
constexpr unsigned XCDS = 8;
__device__ __forceinline__ int64_t xcd_block() {
  return (int64_t)(blockIdx.x % XCDS) * (gridDim.x / XCDS) + blockIdx.x / XCDS;
}
// The (block, grid) a kernel body runs as: the launch's own for a standalone
// launch; one system's share of the grid for a lockstep batch launch.
struct Blk {
  int64_t id, dim;
};

// y = A x, LANES lanes per row, accumulation in T
template <class T>
__global__ void __launch_bounds__(TPB) k_spmv(const int64_t* __restrict__ rp,
                                              const int32_t* __restrict__ col,
                                              const T* __restrict__ val, int64_t n,
                                              const T* __restrict__ x, T* __restrict__ y) {
  const int lane = threadIdx.x % LANES;
  const int64_t groups = (int64_t)gridDim.x * (TPB / LANES);
  for (int64_t row = blockIdx.x * (int64_t)(TPB / LANES) + threadIdx.x / LANES; row < n;
       row += groups) {
    T acc = row_dot<T>(rp, col, val, x, row, lane);
    if (lane == 0) y[row] = acc;
  }
}

// y = diag(scale) (A x), FP64 (one power-iteration step of D^-1 A)
__global__ void __launch_bounds__(TPB) k_spmv_scaled(const int64_t* __restrict__ rp,
                                                     const int32_t* __restrict__ col,
                                                     const double* __restrict__ val, int64_t n,
                                                     const double* __restrict__ scale,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ y) {
  const int lane = threadIdx.x % LANES;
  const int64_t groups = (int64_t)gridDim.x * (TPB / LANES);
  for (int64_t row = blockIdx.x * (int64_t)(TPB / LANES) + threadIdx.x / LANES; row < n;
       row += groups) {
    double acc = row_dot<double>(rp, col, val, x, row, lane);
    if (lane == 0) y[row] = scale[row] * acc;
  }
}

// Partial |y|^2 and |x|^2 of both diagonal blocks in one launch: blocks
// [0, g0) cover rows [0, split) and blocks [g0, g0 + g1) rows [split, n), each
// range strided over its own blocks (the partials of two separate dot
// launches over the ranges).
__global__ void __launch_bounds__(TPB) k_power_sumsq(const double* __restrict__ y,
                                                     const double* __restrict__ x, int64_t split,
                                                     int64_t n, int g0, double* __restrict__ py,
                                                     double* __restrict__ px) {
  __shared__ double sh[TPB / 64];
  const bool b0 = (int)blockIdx.x < g0;
  const int64_t lo = b0 ? 0 : split, hi = b0 ? split : n;
  const int64_t lb = b0 ? blockIdx.x : blockIdx.x - g0;
  const int64_t g = b0 ? g0 : gridDim.x - g0;
  double vy = 0.0, vx = 0.0;
  for (int64_t i = lo + lb * TPB + threadIdx.x; i < hi; i += g * TPB) {
    vy += y[i] * y[i];
    vx += x[i] * x[i];
  }
  const double ty = block_reduce(vy, sh);
  __syncthreads();
  const double tx = block_reduce(vx, sh);
  if (threadIdx.x == 0) {
    py[blockIdx.x] = ty;
    px[blockIdx.x] = tx;
  }
}

// The two blocks' sums in fixed order (block b of the launch: matrix block b):
// sq[2b] = |y_b|^2, sq[2b + 1] = |x_b|^2.
__global__ void __launch_bounds__(TPB) k_power_finish(const double* __restrict__ py,
                                                      const double* __restrict__ px, int g0,
                                                      int g1, double* __restrict__ sq) {
  __shared__ double sh[TPB / 64];
  const int off = blockIdx.x == 0 ? 0 : g0, np = blockIdx.x == 0 ? g0 : g1;
  double vy = 0.0, vx = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    vy += py[off + i];
    vx += px[off + i];
  }
  const double ty = block_reduce(vy, sh);
  __syncthreads();
  const double tx = block_reduce(vx, sh);
  if (threadIdx.x == 0) {
    sq[2 * blockIdx.x] = ty;
    sq[2 * blockIdx.x + 1] = tx;
  }
}

// r = b - A x in FP64 with partial |r|^2
__global__ void __launch_bounds__(TPB) k_residual(const int64_t* __restrict__ rp,
                                                  const int32_t* __restrict__ col,
                                                  const double* __restrict__ val, int64_t n,
                                                  const double* __restrict__ x,
                                                  const double* __restrict__ b,
                                                  double* __restrict__ r, double* part) {
  __shared__ double sh[TPB / 64];
  const int lane = threadIdx.x % LANES;
  const int64_t groups = (int64_t)gridDim.x * (TPB / LANES);
  double loc = 0.0;
  for (int64_t row = blockIdx.x * (int64_t)(TPB / LANES) + threadIdx.x / LANES; row < n;
       row += groups) {
    double acc = row_dot<double>(rp, col, val, x, row, lane);
    if (lane == 0) {
      double ri = b[row] - acc;
      r[row] = ri;
      loc += ri * ri;
    }
  }
  double t = block_reduce(loc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// sum of NBLK partials in a fixed order (one block)
__global__ void k_finish_sum(const double* part, int np, double* out) {
  __shared__ double sh[TPB / 64];
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) v += part[i];
  double t = block_reduce(v, sh);
  if (threadIdx.x == 0) *out = t;
}

template <class T>
__global__ void k_dot(const T* a, const T* b, int64_t n, double* part) {
  __shared__ double sh[TPB / 64];
  double v = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v += to_d(a[i]) * to_d(b[i]);
  double t = block_reduce(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// --- MINRES -----------------------------------------------------------------

// r1 = r2 = b, y = P b, x = w = w2 = 0, partial beta1^2 = b.y
template <class T>
__global__ void k_mr_init(const T* __restrict__ b, const T* __restrict__ dinv, int64_t n,
                          T* r1, T* r2, T* y, T* x, T* w, T* w2, double* part) {
  __shared__ double sh[TPB / 64];
  double v = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    T bi = b[i];
    T yi = bi * dinv[i];
    r1[i] = bi;
    r2[i] = bi;
    y[i] = yi;
    x[i] = 0;
    w[i] = 0;
    w2[i] = 0;
    v += to_d(bi) * to_d(yi);
  }
  double t = block_reduce(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__device__ __forceinline__ void k_mr_start_body(MinresS* S, const double* part, int np, double rtol, long long max_iters, Blk bk) {
  __shared__ double sh[TPB / 64];
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) v += part[i];
  double b2 = block_reduce(v, sh);
  if (threadIdx.x == 0) {
    MinresS s = {};
    s.rtol = rtol;
    s.max_iters = max_iters;
    s.iter = 0;
    if (b2 <= 0.0) {
      s.done = 1;
      s.residual = 0.0;
    } else {
      double beta1 = sqrt(b2);
      s.beta1 = beta1;
      s.beta = beta1;
      s.oldb = 0.0;
      s.dbar = 0.0;
      s.epsln = 0.0;
      s.phibar = beta1;
      s.cs = -1.0;
      s.sn = 0.0;
      s.residual = 1.0;
      s.done = max_iters <= 0;
    }
    *S = s;
  }
}

__global__ void k_mr_start(MinresS* S, const double* part, int np, double rtol, long long max_iters) {
  k_mr_start_body(S, part, np, rtol, max_iters, Blk{blockIdx.x, (int64_t)gridDim.x});
}

// yn = A (y/beta) - (beta/oldb) r1, written over r1; partial alpha = (y/beta).yn
template <class T>
__device__ __forceinline__ void k_mr_spmv_body(const int64_t* __restrict__ rp,
                                                 const int32_t* __restrict__ col,
                                                 const T* __restrict__ val, int64_t n,
                                                 const T* __restrict__ y, T* __restrict__ r1,
                                                 const MinresS* __restrict__ S, double* part, Blk bk) {
  if (S->done) return;
  __shared__ double sh[TPB / 64];
  const T invb = (T)(1.0 / S->beta);
  const T cr = (S->iter >= 1) ? (T)(S->beta / S->oldb) : (T)0;
  const int lane = threadIdx.x % LANES;
  const int64_t groups = (int64_t)bk.dim * (TPB / LANES);
  double loc = 0.0;
  for (int64_t row = bk.id * (int64_t)(TPB / LANES) + threadIdx.x / LANES; row < n;
       row += groups) {
    T acc = row_dot<T>(rp, col, val, y, row, lane);
    if (lane == 0) {
      T yn = acc * invb - cr * r1[row];
      r1[row] = yn;
      loc += to_d(y[row] * invb) * to_d(yn);
    }
  }
  double t = block_reduce(loc, sh);
  if (threadIdx.x == 0) part[bk.id] = t;
}

template <class T>
__global__ void __launch_bounds__(TPB) k_mr_spmv(const int64_t* __restrict__ rp,
                                                 const int32_t* __restrict__ col,
                                                 const T* __restrict__ val, int64_t n,
                                                 const T* __restrict__ y, T* __restrict__ r1,
                                                 const MinresS* __restrict__ S, double* part) {
  k_mr_spmv_body<T>(rp, col, val, n, y, r1, S, part, Blk{blockIdx.x, (int64_t)gridDim.x});
}

// Harmonic border (rank nh): rows/cols n..n+nh-1 of the KKT system.
// partial dots y[0:n] . border_h
template <class T>
__device__ __forceinline__ void k_mr_border_dot_body(const T* __restrict__ border, int nh, int64_t n,
                                const T* __restrict__ y, const MinresS* S, double* part, Blk bk) {
  if (S->done) return;
  __shared__ double sh[TPB / 64];
  for (int h = 0; h < nh; ++h) {
    double v = 0.0;
    for (int64_t i = bk.id * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)bk.dim * blockDim.x)
      v += to_d(border[h * n + i]) * to_d(y[i]);
    double t = block_reduce(v, sh);
    if (threadIdx.x == 0) part[h * NBLK + bk.id] = t;
    __syncthreads();
  }
}

template <class T>
__global__ void k_mr_border_dot(const T* __restrict__ border, int nh, int64_t n,
                                const T* __restrict__ y, const MinresS* S, double* part) {
  k_mr_border_dot_body<T>(border, nh, n, y, S, part, Blk{blockIdx.x, (int64_t)gridDim.x});
}

// border rows of yn and the alpha correction (one block)
template <class T>
__device__ __forceinline__ void k_mr_border_rows_body(int nh, int64_t n, const T* y, T* r1, MinresS* S,
                                 const double* part, int np, Blk bk) {
  if (S->done) return;
  __shared__ double sh[TPB / 64];
  __shared__ double dh[8];
  for (int h = 0; h < nh; ++h) {
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) v += part[h * NBLK + i];
    double t = block_reduce(v, sh);
    if (threadIdx.x == 0) dh[h] = t;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double invb = 1.0 / S->beta;
    double cr = (S->iter >= 1) ? S->beta / S->oldb : 0.0;
    double extra = 0.0;
    for (int h = 0; h < nh; ++h) {
      double vh = to_d(y[n + h]) * invb;
      T ynh = (T)(dh[h] * invb - cr * to_d(r1[n + h]));
      r1[n + h] = ynh;
      extra += vh * (dh[h] * invb) + vh * to_d(ynh);
    }
    S->extra_alpha = extra;
  }
}

template <class T>
__global__ void k_mr_border_rows(int nh, int64_t n, const T* y, T* r1, MinresS* S,
                                 const double* part, int np) {
  k_mr_border_rows_body<T>(nh, n, y, r1, S, part, np, Blk{blockIdx.x, (int64_t)gridDim.x});
}

// yn[0:n] += (y[n+h]/beta) border_h
template <class T>
__device__ __forceinline__ void k_mr_border_axpy_body(const T* __restrict__ border, int nh, int64_t n,
                                 const T* __restrict__ y, T* __restrict__ r1, const MinresS* S, Blk bk) {
  if (S->done) return;
  const T invb = (T)(1.0 / S->beta);
  for (int64_t i = bk.id * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)bk.dim * blockDim.x) {
    T acc = r1[i];
    for (int h = 0; h < nh; ++h) acc += (y[n + h] * invb) * border[h * n + i];
    r1[i] = acc;
  }
}

template <class T>
__global__ void k_mr_border_axpy(const T* __restrict__ border, int nh, int64_t n,
                                 const T* __restrict__ y, T* __restrict__ r1, const MinresS* S) {
  k_mr_border_axpy_body<T>(border, nh, n, y, r1, S, Blk{blockIdx.x, (int64_t)gridDim.x});
}

__device__ __forceinline__ void k_mr_alpha_body(MinresS* S, const double* part, int np, int with_extra, Blk bk) {
  if (S->done) return;
  __shared__ double sh[TPB / 64];
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) v += part[i];
  double t = block_reduce(v, sh);
  if (threadIdx.x == 0) S->alfa = t + (with_extra ? S->extra_alpha : 0.0);
}

__global__ void k_mr_alpha(MinresS* S, const double* part, int np, int with_extra) {
  k_mr_alpha_body(S, part, np, with_extra, Blk{blockIdx.x, (int64_t)gridDim.x});
}

// yn -= (alpha/beta) r2; z = dinv * yn; partial beta^2 = yn.z
template <class T>
__global__ void __launch_bounds__(TPB) k_mr_update(T* __restrict__ yn, const T* __restrict__ r2,
                                                   const T* __restrict__ dinv, T* __restrict__ z,
                                                   int64_t n, const MinresS* __restrict__ S,
                                                   double* part) {
  if (S->done) return;
  __shared__ double sh[TPB / 64];
  const T c = (T)(-S->alfa / S->beta);
  double v = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    T a = yn[i] + c * r2[i];
    T zi = a * dinv[i];
    yn[i] = a;
    z[i] = zi;
    v += to_d(a) * to_d(zi);
  }
  double t = block_reduce(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Givens rotation (krylov.rs:168-186), one thread after the beta^2 reduction.
__device__ __forceinline__ void k_mr_rotate_body(MinresS* S, const double* part, int np, Blk bk) {
  if (S->done) return;
  __shared__ double sh[TPB / 64];
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) v += part[i];
  double b2 = block_reduce(v, sh);
  if (threadIdx.x != 0) return;
  MinresS s = *S;
  const double eps = 2.220446049250313e-16;
  double beta_new = sqrt(fmax(b2, 0.0));
  double alfa = s.alfa;
  s.invb_prev = 1.0 / s.beta;
  s.oldb = s.beta;
  s.beta = beta_new;
  double oldeps = s.epsln;
  double delta = s.cs * s.dbar + s.sn * alfa;
  double gbar = s.sn * s.dbar - s.cs * alfa;
  s.epsln = s.sn * beta_new;
  s.dbar = -s.cs * beta_new;
  double gamma = fmax(sqrt(gbar * gbar + beta_new * beta_new), eps);
  s.cs = gbar / gamma;
  s.sn = beta_new / gamma;
  double phi = s.cs * s.phibar;
  s.phibar *= s.sn;
  s.oldeps = oldeps;
  s.delta = delta;
  s.invg = 1.0 / gamma;
  s.phi = phi;
  s.iter += 1;
  s.residual = s.phibar / s.beta1;
  *S = s;
}

__global__ void k_mr_rotate(MinresS* S, const double* part, int np) {
  k_mr_rotate_body(S, part, np, Blk{blockIdx.x, (int64_t)gridDim.x});
}

// w_new = (y_old/beta_old - oldeps w2 - delta w)/gamma -> into w2's buffer; x += phi w_new.
// Runs in the same iteration as the rotation; the done flag is raised after it.
template <class T>
__device__ __forceinline__ void k_mr_wx_body(const T* __restrict__ yold, T* __restrict__ w,
                                               T* __restrict__ w2, T* __restrict__ x, int64_t n,
                                               MinresS* S, Blk bk) {
  if (S->done) return;
  const T ib = (T)S->invb_prev, oe = (T)S->oldeps, de = (T)S->delta, ig = (T)S->invg,
          ph = (T)S->phi;
  for (int64_t i = bk.id * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)bk.dim * blockDim.x) {
    T wn = (yold[i] * ib - oe * w2[i] - de * w[i]) * ig;
    w2[i] = wn;
    x[i] += ph * wn;
  }
}

template <class T>
__global__ void __launch_bounds__(TPB) k_mr_wx(const T* __restrict__ yold, T* __restrict__ w,
                                               T* __restrict__ w2, T* __restrict__ x, int64_t n,
                                               MinresS* S) {
  k_mr_wx_body<T>(yold, w, w2, x, n, S, Blk{blockIdx.x, (int64_t)gridDim.x});
}

// --- Chebyshev block preconditioner ---------------------------------------
// z = p_m(D^-1 A) D^-1 r per diagonal block of A = blockdiag(hdif_gram(k-1),
// hdif_gram(k)): m steps of the Chebyshev semi-iteration for D^-1 A z = D^-1 r
// from z = 0 on [a_b, b_b] (Saad, Iterative Methods for Sparse Linear Systems,
// Alg. 12.1). The polynomial is fixed and positive on (0, b_b], so the
// preconditioner is SPD, as MINRES needs. coef: [init scale per block] then
// per step k: (rho_{k+1} rho_k, 2 rho_{k+1}/delta) per block. Rows >= n (the
// harmonic border) get the identity.

// yn -= (alpha/beta) r2 (the Lanczos vector update alone)
template <class T>
__global__ void __launch_bounds__(TPB) k_mr_update_only(T* __restrict__ yn, const T* __restrict__ r2,
                                                        int64_t N, const MinresS* __restrict__ S) {
  if (S->done) return;
  const T c = (T)(-S->alfa / S->beta);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x)
    yn[i] = yn[i] + c * r2[i];
}

template <class T>
__global__ void __launch_bounds__(TPB) k_cheb_init(const T* __restrict__ r, const T* __restrict__ dinv,
                                                   T* __restrict__ res, T* __restrict__ d,
                                                   T* __restrict__ z, int64_t n, int64_t N,
                                                   int64_t split, const double* __restrict__ coef,
                                                   const MinresS* __restrict__ S) {
  if (S && S->done) return;
  const T s0 = (T)coef[0], s1 = (T)coef[1];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    T ri = r[i];
    if (i < n) {
      T q = ri * dinv[i];
      res[i] = q;
      d[i] = q * (i < split ? s0 : s1);
      z[i] = 0;
    } else {
      z[i] = ri;
    }
  }
}

// one step: t = (A d)_i; res_i -= dinv_i t; z_i += d_i; dn_i = c1 d_i + c2 res_i.
// UNIT: A is the diagonally scaled D^-1/2 P D^-1/2 (unit diagonal), so the
// Jacobi factor is 1 and dinv is not read.
template <class T, class I, class V = T, bool UNIT = false>
__global__ void __launch_bounds__(TPB) k_cheb_step(const int64_t* __restrict__ rp,
                                                   const I* __restrict__ col,
                                                   const V* __restrict__ val, int64_t n, int64_t split,
                                                   const T* __restrict__ dinv, const T* __restrict__ d,
                                                   T* __restrict__ res, T* __restrict__ z,
                                                   T* __restrict__ dn, const double* __restrict__ coef,
                                                   int step, const MinresS* __restrict__ S) {
  if (S && S->done) return;
  const double* c = coef + 2 + 4 * step;
  const T a0 = (T)c[0], b0 = (T)c[1], a1 = (T)c[2], b1 = (T)c[3];
  const int lane = threadIdx.x % PLANES;
  const int64_t groups = (int64_t)gridDim.x * (TPB / PLANES);
  for (int64_t row = blockIdx.x * (int64_t)(TPB / PLANES) + threadIdx.x / PLANES; row < n;
       row += groups) {
    T acc = row_dot<T, I, V, PLANES>(rp, col, val, d, row, lane);
    if (lane == 0) {
      T di = d[row];
      T ri = UNIT ? res[row] - acc : res[row] - dinv[row] * acc;
      res[row] = ri;
      z[row] += di;
      dn[row] = (row < split) ? (a0 * di + b0 * ri) : (a1 * di + b1 * ri);
    }
  }
}

// z_i += d_i (the last step's direction), times sd_i when the polynomial ran
// on the scaled matrix; partial dot r.z over all N rows
template <class T>
__global__ void __launch_bounds__(TPB) k_cheb_finish(const T* __restrict__ r, const T* __restrict__ d,
                                                     T* __restrict__ z, const T* __restrict__ sd,
                                                     int64_t n, int64_t N,
                                                     const MinresS* __restrict__ S, double* part) {
  if (S && S->done) return;
  __shared__ double sh[TPB / 64];
  double v = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    T zi = z[i];
    if (i < n) {
      zi += d[i];
      if (sd) zi *= sd[i];
      z[i] = zi;
    }
    v += to_d(r[i]) * to_d(zi);
  }
  double t = block_reduce(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ---- the polynomial on the SELL-64 copy of D^-1/2 P D^-1/2 ----
// Three-term form of the same Chebyshev iteration (x_0 = 0, x_1 = f / theta):
//   x_{k+2} = x_{k+1} + a_k (x_{k+1} - x_k) + b_k (f - A x_{k+1}),
// with d_{k+1} = x_{k+2} - x_{k+1} and r_{k+1} = f - A x_{k+1} it is the
// two-term step above, so (a_k, b_k) are the same per-block coefficients. Per
// row it reads x_{k+1} (gathered), x_k and f and writes x_{k+2} over x_k.
// Mixed path: FP32 vectors, entries packed {FP16 value | int16 offset << 16};
// FP64 path: FP64 vectors, FP64 values and int16 offsets in two arrays.

// Packed SELL entry: FP16 value in the low half, int16 (col - row) in the high.
__device__ __forceinline__ float sell_val(uint32_t e) {
  const uint16_t b = (uint16_t)(e & 0xffffu);
  _Float16 h;
  __builtin_memcpy(&h, &b, 2);
  return (float)h;
}
__device__ __forceinline__ int64_t sell_off(uint32_t e) { return (int64_t)(int16_t)(e >> 16); }

constexpr int SELL_UNROLL = 8;
constexpr int SELL_WINDOW = 256;  // rows sorted by length inside windows of this many

// Entry q of the slice storage of row `row` as (value, column). Narrow forms
// keep the column as an int16 offset from the row; the wide forms (when some
// offset does not fit) keep int32 columns in an array of their own.
struct SellF16 {
  const uint32_t* __restrict__ e;
  __device__ __forceinline__ void get(int64_t q, int64_t row, float& v, int64_t& c) const {
    const uint32_t w = e[q];
    v = sell_val(w);
    c = row + sell_off(w);
  }
};
struct SellF64 {
  const double* __restrict__ v;
  const int16_t* __restrict__ o;
  __device__ __forceinline__ void get(int64_t q, int64_t row, double& val, int64_t& c) const {
    val = v[q];
    c = row + o[q];
  }
};
struct SellF16W {
  const _Float16* __restrict__ v;
  const int32_t* __restrict__ col;
  __device__ __forceinline__ void get(int64_t q, int64_t, float& val, int64_t& c) const {
    val = (float)v[q];
    c = col[q];
  }
};
struct SellF64W {
  const double* __restrict__ v;
  const int32_t* __restrict__ col;
  __device__ __forceinline__ void get(int64_t q, int64_t, double& val, int64_t& c) const {
    val = v[q];
    c = col[q];
  }
};

// Off-diagonal part of row `row` (lane row % 64 of its slice) times x; the
// summation order is fixed by the stored pattern.
template <class T, class E>
__device__ __forceinline__ T sell_row(E E_, const int64_t* __restrict__ sp, int64_t row,
                                      const T* __restrict__ x) {
  constexpr int U = SELL_UNROLL;
  const int64_t s = row >> 6;
  const int64_t base = sp[s];
  const int w = (int)((sp[s + 1] - base) >> 6);
  const int64_t p = base + (row & 63);
  T a[U];
#pragma unroll
  for (int u = 0; u < U; ++u) a[u] = 0;
  int j = 0;
  for (; j + U <= w; j += U) {
    T v[U];
    int64_t c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) E_.get(p + (int64_t)(j + u) * 64, row, v[u], c[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] += v[u] * x[c[u]];
  }
  for (; j < w; ++j) {
    T v;
    int64_t c;
    E_.get(p + (int64_t)j * 64, row, v, c);
    a[0] += v * x[c];
  }
  T acc = a[0];
#pragma unroll
  for (int u = 1; u < U; ++u) acc += a[u];
  return acc;
}

// f = sd r, x_0 = 0, x_1 = f / theta_block in the SELL row order (f[i] from
// row perm[i]; the renumbering keeps both blocks in place, so i < split tells
// the block). TP is the polynomial's precision (r is rounded to it first).
template <class T, class TP = T>
__device__ __forceinline__ void k_cheb3_init_body(const T* __restrict__ r,
                                                    const TP* __restrict__ sd,
                                                    const int32_t* __restrict__ perm,
                                                    TP* __restrict__ f, TP* __restrict__ x0,
                                                    TP* __restrict__ x1, int64_t n, int64_t split,
                                                    const double* __restrict__ coef,
                                                    const MinresS* __restrict__ S, Blk bk) {
  if (S && S->done) return;
  const TP s0 = (TP)coef[0], s1 = (TP)coef[1];
  for (int64_t i = bk.id * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)bk.dim * blockDim.x) {
    const int64_t j = perm[i];
    const TP fi = (TP)r[j] * sd[j];
    f[i] = fi;
    x0[i] = 0;
    x1[i] = fi * (i < split ? s0 : s1);
  }
}

template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) k_cheb3_init(const T* __restrict__ r,
                                                    const TP* __restrict__ sd,
                                                    const int32_t* __restrict__ perm,
                                                    TP* __restrict__ f, TP* __restrict__ x0,
                                                    TP* __restrict__ x1, int64_t n, int64_t split,
                                                    const double* __restrict__ coef,
                                                    const MinresS* __restrict__ S) {
  k_cheb3_init_body<T, TP>(r, sd, perm, f, x0, x1, n, split, coef, S, Blk{blockIdx.x, (int64_t)gridDim.x});
}

// The Lanczos update yn += c r2 (c = -alpha/beta) fused with k_cheb3_init:
// lane i updates row perm[i] (a bijection on rows < n), so each row is read
// and written by one lane; the harmonic rows (>= n) are updated in place.
template <class T, class TP = T>
__device__ __forceinline__ void k_mr_update_cheb3_init_body(
    T* __restrict__ yn, const T* __restrict__ r2, const TP* __restrict__ sd,
    const int32_t* __restrict__ perm, TP* __restrict__ f, TP* __restrict__ x0, TP* __restrict__ x1,
    int64_t n, int64_t N, int64_t split, const double* __restrict__ coef,
    const MinresS* __restrict__ S, Blk bk) {
  if (S->done) return;
  const T c = (T)(-S->alfa / S->beta);
  const TP s0 = (TP)coef[0], s1 = (TP)coef[1];
  for (int64_t i = bk.id * (int64_t)blockDim.x + threadIdx.x; i < N;
       i += (int64_t)bk.dim * blockDim.x) {
    const int64_t j = i < n ? perm[i] : i;
    const T a = yn[j] + c * r2[j];
    yn[j] = a;
    if (i < n) {
      const TP fi = (TP)a * sd[j];
      f[i] = fi;
      x0[i] = 0;
      x1[i] = fi * (i < split ? s0 : s1);
    }
  }
}

template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) k_mr_update_cheb3_init(
    T* __restrict__ yn, const T* __restrict__ r2, const TP* __restrict__ sd,
    const int32_t* __restrict__ perm, TP* __restrict__ f, TP* __restrict__ x0, TP* __restrict__ x1,
    int64_t n, int64_t N, int64_t split, const double* __restrict__ coef,
    const MinresS* __restrict__ S) {
  k_mr_update_cheb3_init_body<T, TP>(yn, r2, sd, perm, f, x0, x1, n, N, split, coef, S, Blk{blockIdx.x, (int64_t)gridDim.x});
}


// One lane per row, one block per four 64-row slices: x0 <- x_{k+2}.
template <class T, class E>
__device__ __forceinline__ void k_cheb3_step_body(E entries, const int64_t* __restrict__ sp,
                                                    int64_t n, int64_t split,
                                                    const T* __restrict__ f,
                                                    const T* __restrict__ x1, T* __restrict__ x0,
                                                    const double* __restrict__ coef, int step,
                                                    const MinresS* __restrict__ S, Blk bk) {
  if (S && S->done) return;
  const int64_t row = bk.id * TPB + threadIdx.x;
  if (row >= n) return;
  const double* c = coef + 2 + 4 * step + (row < split ? 0 : 2);
  const T a = (T)c[0], b = (T)c[1];
  const T acc = sell_row<T>(entries, sp, row, x1);
  const T xi = x1[row];
  x0[row] = xi + a * (xi - x0[row]) + b * (f[row] - (xi + acc));
}

template <class T, class E>
__global__ void __launch_bounds__(TPB) k_cheb3_step(E entries, const int64_t* __restrict__ sp,
                                                    int64_t n, int64_t split,
                                                    const T* __restrict__ f,
                                                    const T* __restrict__ x1, T* __restrict__ x0,
                                                    const double* __restrict__ coef, int step,
                                                    const MinresS* __restrict__ S) {
  k_cheb3_step_body<T, E>(entries, sp, n, split, f, x1, x0, coef, step, S, Blk{xcd_block(), (int64_t)gridDim.x});
}

// z = sd x_m on rows < n (x in the SELL row order, precision TP) and z = r on
// the harmonic rows; partial dot r.z
template <class T, class TP = T>
__device__ __forceinline__ void k_cheb3_finish_body(const T* __restrict__ r,
                                                      const TP* __restrict__ x,
                                                      const TP* __restrict__ sd,
                                                      const int32_t* __restrict__ inv,
                                                      T* __restrict__ z, int64_t n, int64_t N,
                                                      const MinresS* __restrict__ S,
                                                      double* part, Blk bk) {
  if (S && S->done) return;
  __shared__ double sh[TPB / 64];
  double v = 0.0;
  for (int64_t i = bk.id * (int64_t)blockDim.x + threadIdx.x; i < N;
       i += (int64_t)bk.dim * blockDim.x) {
    const T ri = r[i];
    const T zi = i < n ? (T)(sd[i] * x[inv[i]]) : ri;
    z[i] = zi;
    v += to_d(ri) * to_d(zi);
  }
  double t = block_reduce(v, sh);
  if (threadIdx.x == 0) part[bk.id] = t;
}

template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) k_cheb3_finish(const T* __restrict__ r,
                                                      const TP* __restrict__ x,
                                                      const TP* __restrict__ sd,
                                                      const int32_t* __restrict__ inv,
                                                      T* __restrict__ z, int64_t n, int64_t N,
                                                      const MinresS* __restrict__ S,
                                                      double* part) {
  k_cheb3_finish_body<T, TP>(r, x, sd, inv, z, n, N, S, part, Blk{blockIdx.x, (int64_t)gridDim.x});
}

template <class T>
__device__ __forceinline__ void k_mr_zero_body(const T* __restrict__ b, int64_t n, T* r1, T* r2, T* x, T* w, T* w2, Blk bk) {
  for (int64_t i = bk.id * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)bk.dim * blockDim.x) {
    T bi = b[i];
    r1[i] = bi;
    r2[i] = bi;
    x[i] = 0;
    w[i] = 0;
    w2[i] = 0;
  }
}

template <class T>
__global__ void k_mr_zero(const T* __restrict__ b, int64_t n, T* r1, T* r2, T* x, T* w, T* w2) {
  k_mr_zero_body<T>(b, n, r1, r2, x, w, w2, Blk{blockIdx.x, (int64_t)gridDim.x});
}

__device__ __forceinline__ void k_mr_check_body(MinresS* S, Blk bk) {
  if (S->done) return;
  if (S->residual <= S->rtol || S->iter >= S->max_iters) S->done = 1;
}

__global__ void k_mr_check(MinresS* S) {
  k_mr_check_body(S, Blk{blockIdx.x, (int64_t)gridDim.x});
}

// Continue a finished MINRES run to a new target: its recurrence and iterate
// are kept, only the stopping test changes.
__global__ void k_mr_resume(MinresS* S, double rtol, long long max_iters) {
  if (threadIdx.x == 0) {
    S->done = 0;
    S->rtol = rtol;
    S->max_iters = max_iters;
  }
}

// --- refinement helpers --------------------------------------------------------
__global__ void k_to_f32_scaled(const double* a, int64_t n, double s, float* b) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)(a[i] * s);
}
__global__ void k_axpy_from_f32(const float* a, int64_t n, double s, double* x) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] += s * (double)a[i];
}
__global__ void k_axpy(const double* a, int64_t n, double s, double* x) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] += s * a[i];
}
__global__ void k_f64_to_f32(const double* a, int64_t n, float* b) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
// border rows of r = b - A x
__global__ void k_border_residual(const double* border, int nh, int64_t n, const double* x,
                                  const double* b, double* r, double* part_extra) {
  __shared__ double sh[TPB / 64];
  for (int h = 0; h < nh; ++h) {
    double v = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) v += border[h * n + i] * x[i];
    double t = block_reduce(v, sh);
    if (threadIdx.x == 0) {
      double rh = b[n + h] - t;
      r[n + h] = rh;
      part_extra[h] = rh * rh;
    }
    __syncthreads();
  }
}
__global__ void k_border_residual_axpy(const double* border, int nh, int64_t n, const double* x,
                                       double* r, double* part) {
  // r[0:n] -= sum_h x[n+h] border_h, then partial |r[0:n]|^2 recomputed
  __shared__ double sh[TPB / 64];
  double v = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double ri = r[i];
    for (int h = 0; h < nh; ++h) ri -= x[n + h] * border[h * n + i];
    r[i] = ri;
    v += ri * ri;
  }
  double t = block_reduce(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void k_jacobi(const double* dd, const double* ii, double* oo, int64_t nn) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < nn) oo[i] = dd ? ii[i] * dd[i] : ii[i];
}
__global__ void k_jacobi32(const float* r, const float* dinv, float* z, int64_t nn) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < nn) z[i] = dinv[i] * r[i];
}
// Power-iteration step of both blocks: with ny = sqrt(sq[2b]) and nx =
// sqrt(sq[2b + 1]) of row i's block b, lambda_b = ny / nx (when nx > 0) and
// x = y / ny (when ny > 0).
__global__ void k_power_normalize(const double* y, int64_t split, int64_t n, const double* sq,
                                  double* x, double* lam) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = i < split ? 0 : 1;
  const double ny = sqrt(sq[2 * b]), nx = sqrt(sq[2 * b + 1]);
  if (i == (b ? split : 0) && nx > 0) lam[b] = ny / nx;
  if (ny > 0) x[i] = (1.0 / ny) * y[i];
}

// The Gershgorin bound of D^-1 P per block: max over the block's rows of
// sum_j |P_ij| dinv_i, an upper bound on lambda_max(D^-1 P). Rows are >= 0, so
// the max runs on the bit patterns (order-preserving for non-negative
// doubles): a wavefront max per block, then one atomic per wavefront.
__global__ void k_gershgorin(const int64_t* rp, const double* val, const double* dinv, int64_t n, int64_t split,
                             unsigned long long* out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double v0 = 0.0, v1 = 0.0;
  if (i < n) {
    double acc = 0.0;
    for (int64_t p = rp[i]; p < rp[i + 1]; ++p) acc += fabs(val[p]);
    acc *= fabs(dinv[i]);
    if (i < split)
      v0 = acc;
    else
      v1 = acc;
  }
  for (int off = 32; off > 0; off >>= 1) {
    v0 = fmax(v0, __shfl_xor(v0, off, 64));
    v1 = fmax(v1, __shfl_xor(v1, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    if (v0 > 0.0) atomicMax(&out[0], (unsigned long long)__double_as_longlong(v0));
    if (v1 > 0.0) atomicMax(&out[1], (unsigned long long)__double_as_longlong(v1));
  }
}

// Chunks of the power-iteration start vector 0.5 + U[0,1) from xorshift64
// (13, 7, 17): thread j steps from the generator's state after j * chunk
// steps, so the vector equals the sequential one bit for bit.
__global__ void k_xorshift_chunks(const unsigned long long* starts, int64_t n, int chunk, double* out) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t i0 = j * chunk;
  if (i0 >= n) return;
  unsigned long long v = starts[j];
  const int64_t i1 = i0 + chunk < n ? i0 + chunk : n;
  for (int64_t i = i0; i < i1; ++i) {
    v ^= v << 13;
    v ^= v >> 7;
    v ^= v << 17;
    out[i] = 0.5 + (double)(v >> 11) * (1.0 / 9007199254740992.0);
  }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + TPB - 1) / TPB); }
inline unsigned red_grid(int64_t n) {
  return (unsigned)std::min<int64_t>(NBLK, std::max<int64_t>(1, (n + TPB - 1) / TPB));
}
// Blocks of a CSR row kernel (grid-stride over rows): one per TPB/lanes rows,
// at most NBLK.
inline unsigned spmv_grid(int64_t n, int lanes = LANES) {
  int64_t rows_per_block = TPB / lanes;
  return (unsigned)std::min<int64_t>(NBLK, std::max<int64_t>(1, (n + rows_per_block - 1) / rows_per_block));
}
inline unsigned cheb_grid(int64_t n) { return spmv_grid(n, PLANES); }
// One lane per row, no grid stride: every block covers four whole 64-row
// slices; rounded up to a multiple of the XCD count for xcd_block().
inline unsigned sell_grid(int64_t n) { return (grid_for(n) + XCDS - 1) / XCDS * XCDS; }

}  // namespace

// ---------------------------------------------------------------------------
// solver workspace
// ---------------------------------------------------------------------------
template <class T>
struct MinresWork {
  DBuf<T> r1, r2, y, z, w, w2, x, b, dinv, border;
  DBuf<T> res, d0, d1;  // Chebyshev work vectors
  DBuf<double> part, part2;
  DBuf<MinresS> S;
  hipGraphExec_t graph = nullptr;
  std::vector<std::shared_ptr<void>> graph_args;  // node arguments of `graph`
  int graph_iters = 0, graph_cheb = -1;
  int64_t n = 0, nh = 0;
  int64_t a_nnz = 0;      // KKT nonzeros (for the segment-length estimate)
  int64_t p_entries = 0;  // stored entries of the preconditioner's matrix
  // block preconditioner (cheb_m == 0: Jacobi with dinv)
  int cheb_m = 0;
  int64_t split = 0;
  const int64_t* prp = nullptr;
  const int32_t* pcol = nullptr;
  const int16_t* pcol16 = nullptr;  // band-offset columns of P, when they fit
  const _Float16* pval16 = nullptr;  // scaled P in FP16 (FP32 path), or null
  const T* sd = nullptr;             // D^-1/2 for the scaled polynomial
  // SELL-64 copy of the scaled P (FP16 entries on the FP32 path, FP64 on the FP64 path)
  using SellN = std::conditional_t<std::is_same_v<T, float>, SellF16, SellF64>;
  using SellW = std::conditional_t<std::is_same_v<T, float>, SellF16W, SellF64W>;
  bool has_sell = false, sell_wide = false;
  SellN sell{};   // int16 column offsets
  SellW sellw{};  // int32 columns
  const int64_t* psell_ptr = nullptr;
  const int32_t* sell_perm = nullptr;
  const int32_t* sell_inv = nullptr;
  const T* pval = nullptr;
  const double* coef = nullptr;
  // FP64 solve whose polynomial runs in FP32 on the FP16 SELL copy
  bool pc32 = false;
  SellF16 sell16{};
  SellF16W sell16w{};
  const float* sd32 = nullptr;
  DBuf<float> res32, d0_32, d1_32;
  // multigrid / auxiliary-space preconditioner (replaces the polynomial)
  Hx* hx = nullptr;
  bool hx_fp32 = false;
  ~MinresWork() {
    if (graph) (void)hipGraphExecDestroy(graph);
  }
};

struct SolverState {
  MinresWork<double> w64;
  MinresWork<float> w32;
  DBuf<double> r, x, part;
  MinresS* host_s = nullptr;  // the context's pinned words
};
static_assert(sizeof(MinresS) <= Ctx::kPinnedBytes, "solver scalars fit the pinned words");

void solver_state_free(SolverState* ws) { delete ws; }

template <class T>
static void minres_alloc(Ctx* ctx, MinresWork<T>& w, int64_t n, int64_t nh) {
  if (w.n == n && w.nh == nh && w.r1.p) return;
  int64_t N = n + nh;
  w.r1.alloc(ctx, N);
  w.r2.alloc(ctx, N);
  w.y.alloc(ctx, N);
  w.z.alloc(ctx, N);
  w.w.alloc(ctx, N);
  w.w2.alloc(ctx, N);
  w.x.alloc(ctx, N);
  w.b.alloc(ctx, N);
  w.dinv.alloc(ctx, N);
  w.border.alloc(ctx, std::max<int64_t>(nh * n, 1));
  w.res.alloc(ctx, N);
  w.d0.alloc(ctx, N);
  w.d1.alloc(ctx, N);
  w.part.alloc(ctx, NBLK * 8);
  w.part2.alloc(ctx, NBLK);
  w.S.alloc(ctx, 1);
  w.n = n;
  w.nh = nh;
  if (w.graph) {
    (void)hipGraphExecDestroy(w.graph);
    w.graph = nullptr;
  }
}

// Point the workspace at the system's preconditioner (matrix values of
// precision T); a captured graph holding other pointers is dropped.
template <class T>
static void bind_preconditioner(KktSystem& s, MinresWork<T>& w, const T* pval,
                                const _Float16* pval16 = nullptr, const T* sd = nullptr,
                                bool sell = false) {
  int m = (s.P.nnz > 0 && s.cheb_m > 0 && s.cheb.p) ? s.cheb_m : 0;
  typename MinresWork<T>::SellN sel{};
  typename MinresWork<T>::SellW selw{};
  const bool wide = sell && s.sell_wide;
  if (sell) {
    if constexpr (std::is_same_v<T, float>) {
      if (wide)
        selw = SellF16W{s.Psell_v16.p, s.Psell_col.p};
      else
        sel = SellF16{s.Psell.p};
    } else {
      if (wide)
        selw = SellF64W{s.Psell64.p, s.Psell_col.p};
      else
        sel = SellF64{s.Psell64.p, s.Psell_off.p};
    }
  }
  const int64_t* psell_ptr = sell ? s.Psell_ptr.p : nullptr;
  const int32_t* sell_perm = sell ? s.sell_perm.p : nullptr;
  const int32_t* sell_inv = sell ? s.sell_inv.p : nullptr;
  if (w.graph && (w.pval != pval || w.coef != s.cheb.p || w.prp != s.P.row_ptr.p ||
                  w.pcol != s.P.col.p || w.pcol16 != s.Pcol16.p || w.pval16 != pval16 ||
                  w.sd != sd || w.cheb_m != m || w.has_sell != sell ||
                  std::memcmp(&w.sell, &sel, sizeof(sel)) != 0 ||
                  std::memcmp(&w.sellw, &selw, sizeof(selw)) != 0 || w.psell_ptr != psell_ptr ||
                  w.sell_perm != sell_perm || w.sell_inv != sell_inv || w.pc32)) {
    (void)hipGraphExecDestroy(w.graph);
    w.graph = nullptr;
  }
  w.pc32 = false;
  w.cheb_m = m;
  w.split = s.split;
  w.prp = s.P.row_ptr.p;
  w.pcol = s.P.col.p;
  w.pcol16 = s.Pcol16.p;
  w.pval = pval;
  w.pval16 = pval16;
  w.sd = sd;
  w.has_sell = sell;
  w.sell_wide = wide;
  w.sell = sel;
  w.sellw = selw;
  w.a_nnz = s.A.nnz;
  w.p_entries = sell ? s.sell_entries : s.P.nnz;
  w.psell_ptr = psell_ptr;
  w.sell_perm = sell_perm;
  w.sell_inv = sell_inv;
  w.coef = s.cheb.p;
}

// One Chebyshev step on the bound storage. SELL copy (three-term): b <- x_{k+2}
// from a = x_{k+1} and b = x_k (z unused). CSR (two-term): direction a, next
// direction into b, z accumulated; scaled mode runs on D^-1/2 P D^-1/2 with
// FP16 values, otherwise on D^-1 P.
template <class T>
static void cheb_step(Enq& q, MinresWork<T>& w, const T* a, T* b, T* z, int k, const MinresS* S) {
  const int64_t n = w.n;
  if (w.has_sell && w.sd) {
    if (w.sell_wide)
      q(k_cheb3_step<T, typename MinresWork<T>::SellW>, sell_grid(n), TPB, w.sellw, w.psell_ptr, n,
        w.split, (const T*)w.res.p, a, b, w.coef, k, S);
    else
      q(k_cheb3_step<T, typename MinresWork<T>::SellN>, sell_grid(n), TPB, w.sell, w.psell_ptr, n,
        w.split, (const T*)w.res.p, a, b, w.coef, k, S);
    return;
  }
  const bool scaled = w.pval16 && w.sd;
  if (scaled && w.pcol16)
    q(k_cheb_step<T, int16_t, _Float16, true>, cheb_grid(n), TPB, w.prp, w.pcol16, w.pval16, n,
      w.split, w.dinv.p, a, w.res.p, z, b, w.coef, k, S);
  else if (scaled)
    q(k_cheb_step<T, int32_t, _Float16, true>, cheb_grid(n), TPB, w.prp, w.pcol, w.pval16, n,
      w.split, w.dinv.p, a, w.res.p, z, b, w.coef, k, S);
  else if (w.pcol16)
    q(k_cheb_step<T, int16_t>, cheb_grid(n), TPB, w.prp, w.pcol16, w.pval, n, w.split, w.dinv.p, a,
      w.res.p, z, b, w.coef, k, S);
  else
    q(k_cheb_step<T, int32_t>, cheb_grid(n), TPB, w.prp, w.pcol, w.pval, n, w.split, w.dinv.p, a,
      w.res.p, z, b, w.coef, k, S);
}

// The polynomial of an FP64 solve run in FP32 on the FP16 SELL copy: the
// mixed solve's kernels and coefficients, with r rounded to FP32 on entry and
// z widened to FP64 on exit (the same values as the mixed solve's preconditioner
// applied to the rounded r).
static void cheb_apply_pc32(Enq& q, MinresWork<double>& w, const double* r, double* z,
                            const MinresS* S, double* part, bool init_done) {
  const int64_t n = w.n, N = w.n + w.nh;
  if (!init_done)
    q(k_cheb3_init<double, float>, red_grid(n), TPB, r, w.sd32, w.sell_perm, w.res32.p,
      w.d0_32.p, w.d1_32.p, n, w.split, w.coef, S);
  float* cur = w.d1_32.p;
  float* prev = w.d0_32.p;
  for (int k = 0; k + 1 < w.cheb_m; ++k) {
    if (w.sell_wide)
      q(k_cheb3_step<float, SellF16W>, sell_grid(n), TPB, w.sell16w, w.psell_ptr, n, w.split,
        (const float*)w.res32.p, (const float*)cur, prev, w.coef, k, S);
    else
      q(k_cheb3_step<float, SellF16>, sell_grid(n), TPB, w.sell16, w.psell_ptr, n, w.split,
        (const float*)w.res32.p, (const float*)cur, prev, w.coef, k, S);
    std::swap(cur, prev);
  }
  q(k_cheb3_finish<double, float>, red_grid(N), TPB, r, (const float*)cur, w.sd32, w.sell_inv, z,
    n, N, S, part);
}

// z = P r with the Chebyshev block preconditioner, plus partial sums of r.z
// into `part` (red_grid(N) blocks).
// With init_done, k_mr_update_cheb3_init has already started the SELL form.
template <class T>
static void cheb_apply(Enq& q, MinresWork<T>& w, const T* r, T* z, const MinresS* S,
                       double* part, bool init_done = false) {
  const int64_t n = w.n, N = w.n + w.nh;
  if constexpr (std::is_same_v<T, double>) {
    if (w.pc32) {
      cheb_apply_pc32(q, w, r, z, S, part, init_done);
      return;
    }
  }
  if (w.has_sell && w.sd) {
    // three-term polynomial on the SELL-64 copy; x_{k+2} overwrites x_k
    if (!init_done)
      q(k_cheb3_init<T>, red_grid(n), TPB, r, w.sd, w.sell_perm, w.res.p, w.d0.p, w.d1.p, n,
        w.split, w.coef, S);
    T* cur = w.d1.p;
    T* prev = w.d0.p;
    for (int k = 0; k + 1 < w.cheb_m; ++k) {
      cheb_step<T>(q, w, cur, prev, nullptr, k, S);
      std::swap(cur, prev);
    }
    q(k_cheb3_finish<T>, red_grid(N), TPB, r, (const T*)cur, w.sd, w.sell_inv, z, n, N, S, part);
    return;
  }
  const bool scaled = w.pval16 && w.sd;
  q(k_cheb_init<T>, red_grid(N), TPB, r, scaled ? w.sd : w.dinv.p, w.res.p, w.d0.p, z, n, N,
    w.split, w.coef, S);
  T* d = w.d0.p;
  T* dn = w.d1.p;
  for (int k = 0; k + 1 < w.cheb_m; ++k) {
    cheb_step<T>(q, w, d, dn, z, k, S);
    std::swap(d, dn);
  }
  q(k_cheb_finish<T>, red_grid(N), TPB, r, d, z, scaled ? w.sd : (const T*)nullptr, n, N, S, part);
}

// Enqueue one MINRES iteration; `par` selects the buffer roles.
template <class T>
static void minres_iteration(Enq& q, MinresWork<T>& w, const int64_t* rp, const int32_t* col,
                             const T* val, int par) {
  const int64_t n = w.n, N = w.n + w.nh;
  T* r1 = par ? w.r2.p : w.r1.p;
  T* r2 = par ? w.r1.p : w.r2.p;
  T* y = par ? w.z.p : w.y.p;
  T* z = par ? w.y.p : w.z.p;
  T* wc = par ? w.w2.p : w.w.p;
  T* w2 = par ? w.w.p : w.w2.p;
  q(k_mr_spmv<T>, spmv_grid(n), TPB, rp, col, val, n, y, r1, w.S.p, w.part.p);
  if (w.nh) {
    q(k_mr_border_dot<T>, red_grid(n), TPB, w.border.p, (int)w.nh, n, y, w.S.p, w.part.p + NBLK);
    q(k_mr_border_rows<T>, 1, TPB, (int)w.nh, n, y, r1, w.S.p, w.part.p + NBLK, (int)red_grid(n));
    q(k_mr_border_axpy<T>, red_grid(n), TPB, w.border.p, (int)w.nh, n, y, r1, w.S.p);
  }
  q(k_mr_alpha, 1, TPB, w.S.p, w.part.p, spmv_grid(n), w.nh ? 1 : 0);
  bool pc32 = false;
  if constexpr (std::is_same_v<T, double>) pc32 = w.pc32 && w.cheb_m > 0;
  if (w.hx) {
    q(k_mr_update_only<T>, red_grid(N), TPB, r1, r2, N, w.S.p);
    if constexpr (std::is_same_v<T, double>)
      hx_enqueue(q, *w.hx, (const double*)r1, z, w.S.p, w.part2.p, w.hx_fp32);
    else
      hx_enqueue(q, *w.hx, (const float*)r1, z, w.S.p, w.part2.p);
  } else if (pc32) {
    if constexpr (std::is_same_v<T, double>) {
      q(k_mr_update_cheb3_init<double, float>, red_grid(N), TPB, r1, r2, w.sd32, w.sell_perm,
        w.res32.p, w.d0_32.p, w.d1_32.p, n, N, w.split, w.coef, w.S.p);
      cheb_apply<T>(q, w, r1, z, w.S.p, w.part2.p, true);
    }
  } else if (w.cheb_m > 0 && w.has_sell && w.sd) {
    q(k_mr_update_cheb3_init<T>, red_grid(N), TPB, r1, r2, w.sd, w.sell_perm, w.res.p, w.d0.p,
      w.d1.p, n, N, w.split, w.coef, w.S.p);
    cheb_apply<T>(q, w, r1, z, w.S.p, w.part2.p, true);
  } else if (w.cheb_m > 0) {
    q(k_mr_update_only<T>, red_grid(N), TPB, r1, r2, N, w.S.p);
    cheb_apply<T>(q, w, r1, z, w.S.p, w.part2.p);
  } else {
    q(k_mr_update<T>, red_grid(N), TPB, r1, r2, w.dinv.p, z, N, w.S.p, w.part2.p);
  }
  q(k_mr_rotate, 1, TPB, w.S.p, w.part2.p, red_grid(N));
  q(k_mr_wx<T>, grid_for(N), TPB, y, wc, w2, w.x.p, N, w.S.p);
  q(k_mr_check, 1, 1, w.S.p);
}

// Iterations per segment. A segment ends with a host read of the MINRES
// scalars, and work that other contexts queue meanwhile waits behind the
// segments already on the device (one request's set-up syncs wait for another
// request's segment), so a segment holds about 16 ms of estimated work: bytes
// per iteration at 4 TB/s plus 4 us per kernel; even, 2..graph_iters.
template <class T>
static int segment_iters(const MinresWork<T>& w, const feec_solve_opts& o) {
  const double vb = sizeof(T), N = (double)(w.n + w.nh);
  double bytes = (vb + 4) * (double)w.a_nnz + 8.0 * (double)(w.n + 1) + 12 * vb * N;
  double launches = 10;
  if (w.cheb_m > 1) {
    const bool sell = w.has_sell || w.pc32;
    const double pb = w.pc32 ? 4.0 : vb;  // bytes per polynomial vector entry
    const double eb = sell ? (pb == 4.0 ? 4.0 : 10.0) + (w.sell_wide ? 2.0 : 0.0)
                           : (w.pval16 ? 4.0 : vb + 4.0);
    bytes += (w.cheb_m - 1) * (eb * (double)w.p_entries + 4 * pb * (double)w.n);
    launches += w.cheb_m + 1;
  }
  if (w.hx) {
    bytes += (double)hx_bytes_per_apply(*w.hx) * (std::is_same_v<T, float> || w.hx_fp32 ? 1.0 : 1.6);
    launches += hx_launches_per_apply(*w.hx);
  }
  const double t = bytes / 4e12 + launches * 4e-6;
  const int seg = (int)std::min(1e6, 0.016 / t);
  const int cap = std::max(2, o.graph_iters - (o.graph_iters & 1));
  return std::clamp(seg - (seg & 1), 2, cap);
}

// Run MINRES on w.b (already filled) into w.x. Returns stats.
// With `resume`, the previous run on w (its recurrence, iterate and iteration
// count) continues to the new rtol for at most max_iters more iterations.
template <class T>
static KrylovStats minres_run(Ctx* ctx, SolverState& ss, MinresWork<T>& w, const int64_t* rp,
                              const int32_t* col, const T* val, double rtol, int64_t max_iters,
                              const feec_solve_opts& o, bool resume = false) {
  hipStream_t st = ctx->stream;
  const int64_t N = w.n + w.nh;
  Enq q{st};
  long long start_iter = 0;
  if (resume) {
    MinresS h{};
    copy_to_host(ctx, &h, w.S.p, sizeof(h));
    start_iter = h.iter;
    q(k_mr_resume, 1, 64, w.S.p, rtol, start_iter + (long long)max_iters);
    // the buffer roles rotate with the executed iterations and a segment
    // starts at parity 0: an odd count takes one iteration outside the graph
    if (start_iter & 1) minres_iteration<T>(q, w, rp, col, val, 1);
  } else if (w.hx) {
    q(k_mr_zero<T>, red_grid(N), TPB, w.b.p, N, w.r1.p, w.r2.p, w.x.p, w.w.p, w.w2.p);
    if constexpr (std::is_same_v<T, double>)
      hx_enqueue(q, *w.hx, (const double*)w.b.p, w.y.p, nullptr, w.part.p, w.hx_fp32);
    else
      hx_enqueue(q, *w.hx, (const float*)w.b.p, w.y.p, nullptr, w.part.p);
  } else if (w.cheb_m > 0) {
    q(k_mr_zero<T>, red_grid(N), TPB, w.b.p, N, w.r1.p, w.r2.p, w.x.p, w.w.p, w.w2.p);
    cheb_apply<T>(q, w, w.b.p, w.y.p, nullptr, w.part.p);
  } else {
    q(k_mr_init<T>, red_grid(N), TPB, w.b.p, w.dinv.p, N, w.r1.p, w.r2.p, w.y.p, w.x.p, w.w.p,
      w.w2.p, w.part.p);
  }
  if (!resume) q(k_mr_start, 1, TPB, w.S.p, w.part.p, red_grid(N), rtol, (long long)max_iters);
  const int seg = segment_iters(w, o);
  if (o.use_graph && (!w.graph || w.graph_iters != seg || w.graph_cheb != (w.hx ? -2 : w.cheb_m))) {
    if (w.graph) {
      (void)hipGraphExecDestroy(w.graph);
      w.graph = nullptr;
    }
    // Build the segment graph node by node; it is destroyed on any failure.
    hipGraph_t g = nullptr;
    FEEC_HIP(hipGraphCreate(&g, 0));
    w.graph_args.clear();
    Enq gq{st, g, &w.graph_args};
    try {
      for (int i = 0; i < seg; ++i) minres_iteration<T>(gq, w, rp, col, val, i & 1);
    } catch (...) {
      (void)hipGraphDestroy(g);
      throw;
    }
    hipError_t ie = hipGraphInstantiate(&w.graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    FEEC_HIP(ie);
    w.graph_iters = seg;
    w.graph_cheb = w.hx ? -2 : w.cheb_m;
  }
  KrylovStats ks;
  // One segment at a time; the host reads the scalar block between segments.
  while (true) {
    if (o.use_graph) {
      FEEC_HIP(hipGraphLaunch(w.graph, st));
    } else {
      for (int i = 0; i < seg; ++i) minres_iteration<T>(q, w, rp, col, val, i & 1);
    }
    FEEC_HIP(hipMemcpyAsync(ss.host_s, w.S.p, sizeof(MinresS), hipMemcpyDeviceToHost, st));
    FEEC_HIP(hipStreamSynchronize(st));
    if (ss.host_s->done) break;
  }
  ks.iters = ss.host_s->iter - start_iter;
  ks.precond_residual = ss.host_s->residual;
  ks.converged = ss.host_s->residual <= rtol;
  return ks;
}

// Scaled preconditioner matrix D^-1/2 P D^-1/2 in FP16 (unit diagonal,
// entries in [-1, 1] for an SPD P) and D^-1/2 in FP32.
__global__ void k_scale_half(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                             const double* __restrict__ val, const double* __restrict__ inv_diag,
                             int64_t n, _Float16* __restrict__ out, float* __restrict__ sd) {
  for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < n;
       row += (int64_t)gridDim.x * blockDim.x) {
    const double si = sqrt(fabs(inv_diag[row]));
    sd[row] = (float)si;
    for (int64_t p = rp[row]; p < rp[row + 1]; ++p)
      out[p] = (_Float16)(val[p] * si * sqrt(fabs(inv_diag[col[p]])));
  }
}

// Off-diagonal entries per row.
__global__ void k_row_offdiag(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                              int64_t n, int32_t* __restrict__ cnt) {
  for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < n;
       row += (int64_t)gridDim.x * blockDim.x) {
    int32_t c = 0;
    for (int64_t p = rp[row]; p < rp[row + 1]; ++p) c += col[p] != row;
    cnt[row] = c;
  }
}

// One block per window of SELL_WINDOW rows (windows restart at `split`): the
// window's rows sorted by off-diagonal count, longest first, ties by row.
__global__ void __launch_bounds__(SELL_WINDOW) k_sell_rank(const int32_t* __restrict__ cnt,
                                                           int64_t n, int64_t split,
                                                           int32_t* __restrict__ perm,
                                                           int32_t* __restrict__ inv) {
  __shared__ int32_t sc[SELL_WINDOW];
  const int64_t nw0 = (split + SELL_WINDOW - 1) / SELL_WINDOW;
  const int64_t b = blockIdx.x;
  const int64_t lo = b < nw0 ? b * SELL_WINDOW : split + (b - nw0) * SELL_WINDOW;
  const int64_t hi = b < nw0 ? (lo + SELL_WINDOW < split ? lo + SELL_WINDOW : split)
                             : (lo + SELL_WINDOW < n ? lo + SELL_WINDOW : n);
  const int t = threadIdx.x;
  const int m = (int)(hi - lo);
  const int32_t c = t < m ? cnt[lo + t] : 0;
  sc[t] = c;
  __syncthreads();
  if (t >= m) return;
  int rank = 0;
  for (int u = 0; u < m; ++u) rank += (sc[u] > c) || (sc[u] == c && u < t);
  perm[lo + rank] = (int32_t)(lo + t);
  inv[lo + t] = (int32_t)(lo + rank);
}

// Longest row of each 64-row slice in the SELL order, and the sum of all
// counts into *total.
__global__ void k_sell_width(const int32_t* __restrict__ cnt, const int32_t* __restrict__ perm,
                             int64_t n, int64_t nslices, int64_t* __restrict__ width,
                             unsigned long long* total) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= nslices) return;
  int64_t w = 0, sum = 0;
  const int64_t r1 = (s + 1) * 64 < n ? (s + 1) * 64 : n;
  for (int64_t i = s * 64; i < r1; ++i) {
    const int64_t c = cnt[perm[i]];
    w = c > w ? c : w;
    sum += c;
  }
  width[s] = w;
  if (sum) atomicAdd(total, (unsigned long long)sum);
}

// Largest |inv[col] - inv[row]| over the off-diagonal entries.
__global__ void k_sell_band(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                            const int32_t* __restrict__ inv, int64_t n,
                            unsigned long long* maxdist) {
  unsigned long long md = 0;
  for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < n;
       row += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ir = inv[row];
    for (int64_t p = rp[row]; p < rp[row + 1]; ++p) {
      const int64_t dlt = (int64_t)inv[col[p]] - ir;
      const unsigned long long a = (unsigned long long)(dlt < 0 ? -dlt : dlt);
      md = a > md ? a : md;
    }
  }
  if (md) atomicMax(maxdist, md);
}

// Scaled off-diagonal values of row perm[i] into lane i of the SELL copy, in
// P's column order, with zero entries (value 0, offset 0) after the last one.
__global__ void k_sell_fill(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                            const double* __restrict__ val, const double* __restrict__ inv_diag,
                            const int64_t* __restrict__ sp, const int32_t* __restrict__ perm,
                            const int32_t* __restrict__ inv, int64_t n, int64_t nrows_padded,
                            uint32_t* __restrict__ E) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nrows_padded;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = i >> 6, base = sp[s] + (i & 63);
    const int64_t w = (sp[s + 1] - sp[s]) >> 6;
    int64_t j = 0;
    if (i < n) {
      const int64_t row = perm[i];
      const double si = sqrt(fabs(inv_diag[row]));
      for (int64_t p = rp[row]; p < rp[row + 1]; ++p) {
        const int64_t c = col[p];
        if (c == row) continue;
        const _Float16 h = (_Float16)(val[p] * si * sqrt(fabs(inv_diag[c])));
        uint16_t hb;
        __builtin_memcpy(&hb, &h, 2);
        const uint16_t ob = (uint16_t)(int16_t)(inv[c] - i);
        E[base + j * 64] = (uint32_t)hb | ((uint32_t)ob << 16);
        ++j;
      }
    }
    for (; j < w; ++j) E[base + j * 64] = 0u;
  }
}

// Scaled off-diagonal values of row perm[i] into lane i of the slices (FP16
// or FP64), in P's column order, zero after the row's last entry.
template <class V>
__global__ void k_sell_vals(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                            const double* __restrict__ val, const double* __restrict__ inv_diag,
                            const int64_t* __restrict__ sp, const int32_t* __restrict__ perm,
                            int64_t n, int64_t nrows_padded, V* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nrows_padded;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = i >> 6, base = sp[s] + (i & 63);
    const int64_t w = (sp[s + 1] - sp[s]) >> 6;
    int64_t j = 0;
    if (i < n) {
      const int64_t row = perm[i];
      const double si = sqrt(fabs(inv_diag[row]));
      for (int64_t p = rp[row]; p < rp[row + 1]; ++p) {
        const int64_t c = col[p];
        if (c == row) continue;
        out[base + j * 64] = (V)(val[p] * si * sqrt(fabs(inv_diag[c])));
        ++j;
      }
    }
    for (; j < w; ++j) out[base + j * 64] = (V)0;
  }
}

// Columns of the same entries in the SELL numbering: int16 offsets
// inv[col] - i (REL) or int32 columns inv[col]; padding 0 / the row itself.
template <class I, bool REL>
__global__ void k_sell_cols(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                            const int64_t* __restrict__ sp, const int32_t* __restrict__ perm,
                            const int32_t* __restrict__ inv, int64_t n, int64_t nrows_padded,
                            I* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nrows_padded;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = i >> 6, base = sp[s] + (i & 63);
    const int64_t w = (sp[s + 1] - sp[s]) >> 6;
    int64_t j = 0;
    if (i < n) {
      const int64_t row = perm[i];
      for (int64_t p = rp[row]; p < rp[row + 1]; ++p) {
        const int64_t c = col[p];
        if (c == row) continue;
        out[base + j * 64] = REL ? (I)(inv[c] - i) : (I)inv[c];
        ++j;
      }
    }
    const I pad = REL ? (I)0 : (I)(i < n ? i : 0);
    for (; j < w; ++j) out[base + j * 64] = pad;
  }
}

__global__ void k_scale_sd64(const double* __restrict__ inv_diag, int64_t n, double* __restrict__ sd) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    sd[i] = sqrt(fabs(inv_diag[i]));
}

// D^-1/2 in FP32
__global__ void k_scale_sd(const double* __restrict__ inv_diag, int64_t n, float* __restrict__ sd) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    sd[i] = (float)sqrt(fabs(inv_diag[i]));
}

static void ensure_band(KktSystem& s);

// FP16 CSR copy of D^-1/2 P D^-1/2 (and D^-1/2) from P's FP64 values.
static void kkt_refresh_half_csr(KktSystem& s) {
  if (!s.P16.p || (int64_t)s.P16.n != s.P.nnz) s.P16.alloc(s.ctx, s.P.nnz);
  k_scale_half<<<grid_for(s.n), TPB, 0, s.ctx->stream>>>(s.P.row_ptr.p, s.P.col.p, s.P.val.p,
                                                         s.inv_diag.p, s.n, s.P16.p, s.sdinv32.p);
}

void kkt_refresh_fp32(KktSystem& s) {
  if (!s.val32.p || (int64_t)s.val32.n != s.A.nnz) s.val32.alloc(s.ctx, s.A.nnz);
  if (s.A.nnz)
    k_f64_to_f32<<<grid_for(s.A.nnz), TPB, 0, s.ctx->stream>>>(s.A.val.p, s.A.nnz, s.val32.p);
  int64_t N = s.n + s.nh;
  if (!s.inv_diag32.p || (int64_t)s.inv_diag32.n != N) s.inv_diag32.alloc(s.ctx, N);
  k_f64_to_f32<<<grid_for(N), TPB, 0, s.ctx->stream>>>(s.inv_diag.p, N, s.inv_diag32.p);
  // the polynomial's copies of P are not needed when the multigrid
  // preconditioner replaces it
  if (s.P.nnz && !s.hx) {
    hipStream_t st = s.ctx->stream;
    ensure_band(s);
    if (!s.P32.p || (int64_t)s.P32.n != s.P.nnz) s.P32.alloc(s.ctx, s.P.nnz);
    k_f64_to_f32<<<grid_for(s.P.nnz), TPB, 0, st>>>(s.P.val.p, s.P.nnz, s.P32.p);
    if (!s.sdinv32.p || (int64_t)s.sdinv32.n != s.n) s.sdinv32.alloc(s.ctx, std::max<int64_t>(s.n, 1));
    k_scale_sd<<<grid_for(s.n), TPB, 0, st>>>(s.inv_diag.p, s.n, s.sdinv32.p);
    // the FP16 CSR copy is the fallback when there is no SELL copy, or kept
    // up to date once a solve has asked for it
    if (!s.Psell_ptr.p || s.P16.p) kkt_refresh_half_csr(s);
    if (s.Psell_ptr.p) {
      const int64_t E = std::max<int64_t>(s.sell_entries, 1);
      const int64_t padded = (s.n + 63) / 64 * 64;
      if (!s.sell_wide) {
        if (!s.Psell.p || (int64_t)s.Psell.n != E) s.Psell.alloc(s.ctx, E);
        k_sell_fill<<<grid_for(padded), TPB, 0, st>>>(s.P.row_ptr.p, s.P.col.p, s.P.val.p,
                                                      s.inv_diag.p, s.Psell_ptr.p, s.sell_perm.p,
                                                      s.sell_inv.p, s.n, padded, s.Psell.p);
      } else {
        if (!s.Psell_v16.p || (int64_t)s.Psell_v16.n != E) s.Psell_v16.alloc(s.ctx, E);
        k_sell_vals<_Float16><<<grid_for(padded), TPB, 0, st>>>(
            s.P.row_ptr.p, s.P.col.p, s.P.val.p, s.inv_diag.p, s.Psell_ptr.p, s.sell_perm.p, s.n,
            padded, s.Psell_v16.p);
      }
    }
  }
  FEEC_HIP(hipGetLastError());
}

void kkt_refresh_sell64(KktSystem& s) {
  if (!s.P.nnz) return;
  ensure_band(s);
  if (!s.Psell_ptr.p) return;
  hipStream_t st = s.ctx->stream;
  const int64_t n = s.n, E = std::max<int64_t>(s.sell_entries, 1);
  const int64_t padded = (n + 63) / 64 * 64;
  if (!s.Psell64.p || (int64_t)s.Psell64.n != E) s.Psell64.alloc(s.ctx, E);
  if (!s.sell_wide && !s.Psell_off.p) {  // pattern-level, once
    s.Psell_off.alloc(s.ctx, E);
    k_sell_cols<int16_t, true><<<grid_for(padded), TPB, 0, st>>>(
        s.P.row_ptr.p, s.P.col.p, s.Psell_ptr.p, s.sell_perm.p, s.sell_inv.p, n, padded,
        s.Psell_off.p);
  }
  if (!s.sdinv64.p || (int64_t)s.sdinv64.n != std::max<int64_t>(n, 1))
    s.sdinv64.alloc(s.ctx, std::max<int64_t>(n, 1));
  k_scale_sd64<<<grid_for(n), TPB, 0, st>>>(s.inv_diag.p, n, s.sdinv64.p);
  k_sell_vals<double><<<grid_for(padded), TPB, 0, st>>>(s.P.row_ptr.p, s.P.col.p, s.P.val.p,
                                                        s.inv_diag.p, s.Psell_ptr.p, s.sell_perm.p,
                                                        n, padded, s.Psell64.p);
  FEEC_HIP(hipGetLastError());
}

// Power iteration on D^-1 P, one estimate per diagonal block; then the
// Chebyshev interval [b lo, b] with b = 1.05 lambda_max and lo chosen by degree.
// Band-offset copy of P's columns: col - row as int16 (one pass), and the
// largest |col - row| so the caller keeps it only when it fits.
__global__ void k_band16(const int64_t* __restrict__ rp, const int32_t* __restrict__ col, int64_t n,
                         int16_t* __restrict__ off, unsigned long long* maxdist) {
  unsigned long long md = 0;
  for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < n;
       row += (int64_t)gridDim.x * blockDim.x)
    for (int64_t p = rp[row]; p < rp[row + 1]; ++p) {
      const int64_t dlt = (int64_t)col[p] - row;
      const unsigned long long a = (unsigned long long)(dlt < 0 ? -dlt : dlt);
      md = a > md ? a : md;
      off[p] = (int16_t)(a <= 32767 ? dlt : 0);
    }
  if (md) atomicMax(maxdist, md);
}

// Pattern-level set-up of P, once: int16 band-offset columns when every
// |col - row| <= 32767, the SELL-64 row order and slice offsets, and the
// slices' int32 columns when a renumbered offset does not fit int16.
static void ensure_band(KktSystem& s) {
  if (s.band_checked || s.P.nnz == 0) return;
  Ctx* ctx = s.ctx;
  hipStream_t st = ctx->stream;
  s.band_checked = true;
  s.Pcol16.alloc(ctx, s.P.nnz);
  FEEC_HIP(hipMemsetAsync(ctx->dcount, 0, sizeof(unsigned long long), st));
  k_band16<<<grid_for(s.n), TPB, 0, st>>>(s.P.row_ptr.p, s.P.col.p, s.n, s.Pcol16.p, ctx->dcount);
  unsigned long long md = 0;
  FEEC_HIP(hipMemcpyAsync(&md, ctx->dcount, sizeof(md), hipMemcpyDeviceToHost, st));
  ctx->sync();
  if (md > 32767) s.Pcol16.reset();
  // SELL row order: rows sorted by length inside windows, then its band
  const int64_t n = s.n, ns = (n + 63) / 64;
  DBuf<int32_t> cnt(ctx, n);
  s.sell_perm.alloc(ctx, n);
  s.sell_inv.alloc(ctx, n);
  k_row_offdiag<<<grid_for(n), TPB, 0, st>>>(s.P.row_ptr.p, s.P.col.p, n, cnt.p);
  const int64_t nwin = (s.split + SELL_WINDOW - 1) / SELL_WINDOW +
                       (n - s.split + SELL_WINDOW - 1) / SELL_WINDOW;
  if (nwin > 0)
    k_sell_rank<<<(unsigned)nwin, SELL_WINDOW, 0, st>>>(cnt.p, n, s.split, s.sell_perm.p,
                                                        s.sell_inv.p);
  FEEC_HIP(hipMemsetAsync(ctx->dcount, 0, sizeof(unsigned long long), st));
  k_sell_band<<<grid_for(n), TPB, 0, st>>>(s.P.row_ptr.p, s.P.col.p, s.sell_inv.p, n, ctx->dcount);
  FEEC_HIP(hipMemcpyAsync(&md, ctx->dcount, sizeof(md), hipMemcpyDeviceToHost, st));
  ctx->sync();
  s.sell_wide = md > 32767;
  DBuf<int64_t> width(ctx, ns);
  FEEC_HIP(hipMemsetAsync(ctx->dcount, 0, sizeof(unsigned long long), st));
  k_sell_width<<<grid_for(ns), TPB, 0, st>>>(cnt.p, s.sell_perm.p, n, ns, width.p, ctx->dcount);
  std::vector<int64_t> w(ns), ptr(ns + 1, 0);
  unsigned long long offdiag = 0;
  FEEC_HIP(hipMemcpyAsync(w.data(), width.p, ns * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  FEEC_HIP(hipMemcpyAsync(&offdiag, ctx->dcount, sizeof(offdiag), hipMemcpyDeviceToHost, st));
  ctx->sync();
  for (int64_t i = 0; i < ns; ++i) ptr[i + 1] = ptr[i] + 64 * w[i];
  s.Psell_ptr.alloc(ctx, ns + 1);
  FEEC_HIP(hipMemcpyAsync(s.Psell_ptr.p, ptr.data(), (ns + 1) * sizeof(int64_t),
                          hipMemcpyHostToDevice, st));
  ctx->sync();
  s.sell_entries = ptr[ns];
  s.P_offdiag = (int64_t)offdiag;
  if (s.sell_wide) {  // int32 columns, shared by both precisions
    const int64_t padded = ns * 64;
    s.Psell_col.alloc(ctx, std::max<int64_t>(s.sell_entries, 1));
    k_sell_cols<int32_t, false><<<grid_for(padded), TPB, 0, st>>>(
        s.P.row_ptr.p, s.P.col.p, s.Psell_ptr.p, s.sell_perm.p, s.sell_inv.p, n, padded,
        s.Psell_col.p);
    FEEC_HIP(hipGetLastError());
  }
}

// The start vector of the spectrum estimates: x_i = 0.5 + U[0,1) from the
// (i+1)-th state of xorshift64 (13, 7, 17) seeded 0x2545F4914F6CDD1D. The
// generator is linear over GF(2), so the state after k steps is T^k applied to
// the seed, T the 64x64 bit matrix of one step; the host steps through the
// chunk starts with T^chunk and the device expands the chunks.
void xorshift_start(Ctx* ctx, double* out, int64_t n) {
  if (n <= 0) return;
  constexpr int kChunk = 256;
  using Mat = std::array<uint64_t, 64>;  // column b = the image of bit b
  auto step = [](uint64_t v) {
    v ^= v << 13;
    v ^= v >> 7;
    v ^= v << 17;
    return v;
  };
  auto apply = [](const Mat& m, uint64_t v) {
    uint64_t r = 0;
    for (int b = 0; b < 64; ++b)
      if ((v >> b) & 1) r ^= m[b];
    return r;
  };
  Mat t, p;
  for (int b = 0; b < 64; ++b) t[b] = step(uint64_t(1) << b);
  // p = T^kChunk by squaring (kChunk is a power of two)
  p = t;
  for (int e = 1; e < kChunk; e *= 2) {
    Mat q;
    for (int b = 0; b < 64; ++b) q[b] = apply(p, p[b]);
    p = q;
  }
  const int64_t nchunks = (n + kChunk - 1) / kChunk;
  std::vector<unsigned long long> starts(nchunks);
  uint64_t v = 0x2545F4914F6CDD1Dull;
  for (int64_t j = 0; j < nchunks; ++j) {
    starts[j] = v;
    v = apply(p, v);
  }
  DBuf<unsigned long long> d(ctx, nchunks);
  copy_to_device(ctx, d.p, starts.data(), nchunks * 8);
  k_xorshift_chunks<<<grid_for(nchunks), TPB, 0, ctx->stream>>>(d.p, n, kChunk, out);
  FEEC_HIP(hipGetLastError());
  ctx->sync();
}

constexpr double kChebLmaxMargin = 1.05;
constexpr double kChebGershgorinTight = 1.1;

void kkt_set_chebyshev(KktSystem& s, int m, double lower) {
  Ctx* ctx = s.ctx;
  hipStream_t st = ctx->stream;
  s.cheb_m = (s.P.nnz > 0) ? std::max(m, 0) : 0;
  if (s.cheb_m == 0) return;
  ensure_band(s);
  const int64_t n = s.n;
  DBuf<double> x(ctx, n), y(ctx, n), part(ctx, NBLK + 1);
  xorshift_start(ctx, x.p, n);
  // per block: |y|^2 and |x|^2 (fixed-order sums) and lambda, all on the device
  // so the 30 iterations run without a host round trip
  // four launches per iteration: y = D^-1 P x; per block |y|^2, |x|^2 (fixed
  // order); lambda_b and x = y / |y| per block
  DBuf<double> sq(ctx, 4), dlam(ctx, 2), py(ctx, 2 * NBLK), px(ctx, 2 * NBLK);
  double lam[2] = {1.0, 1.0};
  copy_to_device(ctx, dlam.p, lam, sizeof(lam));
  const int g0 = s.split > 0 ? (int)red_grid(s.split) : 0;
  const int g1 = n > s.split ? (int)red_grid(n - s.split) : 0;
  for (int it = 0; it < 30; ++it) {
    k_spmv_scaled<<<spmv_grid(n), TPB, 0, st>>>(s.P.row_ptr.p, s.P.col.p, s.P.val.p, n,
                                                s.inv_diag.p, x.p, y.p);
    k_power_sumsq<<<g0 + g1, TPB, 0, st>>>(y.p, x.p, s.split, n, g0, py.p, px.p);
    k_power_finish<<<2, TPB, 0, st>>>(py.p, px.p, g0, g1, sq.p);
    k_power_normalize<<<grid_for(n), TPB, 0, st>>>(y.p, s.split, n, sq.p, x.p, dlam.p);
  }
  DBuf<unsigned long long> gb(ctx, 2);
  k_zero_u64<<<1, 64, 0, st>>>(gb.p, 2);
  k_gershgorin<<<grid_for(n), TPB, 0, st>>>(s.P.row_ptr.p, s.P.val.p, s.inv_diag.p, n, s.split, gb.p);
  FEEC_HIP(hipGetLastError());
  copy_to_host(ctx, lam, dlam.p, sizeof(lam));
  unsigned long long gbits[2] = {0, 0};
  copy_to_host(ctx, gbits, gb.p, sizeof(gbits));
  // interval lower ends (fraction of b) tuned on 3D k=1 Kuhn meshes
  const double frac =
      lower > 0.0 ? lower
      : m <= 4    ? 1e-2
      : m <= 8    ? 3e-3
      : m <= 12   ? 1.5e-3
      : m <= 32   ? 1e-3
      : m <= 48   ? 5e-4
      : m <= 64   ? 2.5e-4
      : m <= 96   ? 1.2e-4
                  : 6e-5;
  std::vector<double> coef(2 + 4 * std::max(m - 1, 0), 0.0);
  double rho[2], sigma[2], delta[2];
  for (int b = 0; b < 2; ++b) {
    s.lambda_max[b] = lam[b];
    // the interval's top: the Gershgorin bound where it is within 10% of the
    // power estimate (it bounds lambda_max, which 30 power steps can
    // underestimate by more than 5% when the top of the spectrum is
    // clustered), else 1.05 times the estimate
    double gersh;
    std::memcpy(&gersh, &gbits[b], sizeof(double));
    double bb = (gersh > 0.0 && gersh <= kChebGershgorinTight * lam[b]) ? gersh : kChebLmaxMargin * lam[b];
    double aa = bb * frac;
    double theta = 0.5 * (bb + aa);
    delta[b] = 0.5 * (bb - aa);
    sigma[b] = theta / delta[b];
    rho[b] = 1.0 / sigma[b];
    coef[b] = 1.0 / theta;
  }
  for (int k = 0; k + 1 < m; ++k)
    for (int b = 0; b < 2; ++b) {
      double rn = 1.0 / (2.0 * sigma[b] - rho[b]);
      coef[2 + 4 * k + 2 * b] = rn * rho[b];
      coef[2 + 4 * k + 2 * b + 1] = 2.0 * rn / delta[b];
      rho[b] = rn;
    }
  if ((int64_t)s.cheb.n != (int64_t)coef.size()) s.cheb.alloc(ctx, coef.size());
  FEEC_HIP(hipMemcpyAsync(s.cheb.p, coef.data(), coef.size() * 8, hipMemcpyHostToDevice, st));
  ctx->sync();
}

// |b - A x| / |b| with the FP64 KKT operator (CSR + border); r into ss.r
static double true_residual(KktSystem& s, SolverState& ss, const double* b, const double* x,
                            double bnorm) {
  Ctx* ctx = s.ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = s.n, nh = s.nh;
  k_residual<<<spmv_grid(n), TPB, 0, st>>>(s.A.row_ptr.p, s.A.col.p, s.A.val.p, n, x, b, ss.r.p,
                                           ss.part.p);
  double* fin = ss.part.p + NBLK * 4;
  if (nh) {
    k_border_residual<<<1, TPB, 0, st>>>(s.border.p, (int)nh, n, x, b, ss.r.p, ss.part.p + NBLK * 2);
    k_border_residual_axpy<<<red_grid(n), TPB, 0, st>>>(s.border.p, (int)nh, n, x, ss.r.p, ss.part.p);
    k_finish_sum<<<1, TPB, 0, st>>>(ss.part.p, red_grid(n), fin);
    k_finish_sum<<<1, TPB, 0, st>>>(ss.part.p + NBLK * 2, (int)nh, fin + 1);
  } else {
    k_finish_sum<<<1, TPB, 0, st>>>(ss.part.p, spmv_grid(n), fin);
  }
  double h[2] = {0.0, 0.0};
  FEEC_HIP(hipMemcpyAsync(h, fin, (nh ? 2 : 1) * sizeof(double), hipMemcpyDeviceToHost, st));
  ctx->sync();
  return std::sqrt(h[0] + h[1]) / bnorm;
}

static double device_norm(Ctx* ctx, SolverState& ss, const double* a, int64_t n) {
  k_dot<double><<<red_grid(n), TPB, 0, ctx->stream>>>(a, a, n, ss.part.p);
  double* fin = ss.part.p + NBLK * 4;
  k_finish_sum<<<1, TPB, 0, ctx->stream>>>(ss.part.p, red_grid(n), fin);
  double h = 0.0;
  FEEC_HIP(hipMemcpyAsync(&h, fin, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  return std::sqrt(h);
}

// The polynomial's matrix for each path and feec_pc_storage (the SELL copy
// falls back to CSR where the pattern allows none), made on first use.
// The multigrid / auxiliary-space preconditioner in place of the polynomial.
template <class T>
static void bind_hx(KktSystem& s, MinresWork<T>& w, bool fp32) {
  if (w.graph && (w.hx != s.hx || w.hx_fp32 != fp32 || w.cheb_m != 0)) {
    (void)hipGraphExecDestroy(w.graph);
    w.graph = nullptr;
  }
  w.hx = s.hx;
  w.hx_fp32 = fp32;
  w.pc32 = false;
  w.cheb_m = 0;
  w.has_sell = false;
  w.split = s.split;
  w.a_nnz = s.A.nnz;
  w.p_entries = 0;
}

static void bind_fp64(KktSystem& s, MinresWork<double>& w, int storage) {
  if (s.hx) {
    bind_hx(s, w, false);
    return;
  }
  w.hx = nullptr;
  if (storage == FEEC_PC_STORE_SELL && s.P.nnz && !s.Psell64.p) kkt_refresh_sell64(s);
  if (storage == FEEC_PC_STORE_SELL && s.Psell64.p)
    bind_preconditioner(s, w, s.P.val.p, nullptr, s.sdinv64.p, true);
  else
    bind_preconditioner(s, w, s.P.val.p);
}

static void bind_fp32(KktSystem& s, MinresWork<float>& w, int storage) {
  if (!s.val32.p) kkt_refresh_fp32(s);
  if (s.hx) {
    bind_hx(s, w, true);
    return;
  }
  w.hx = nullptr;
  if (storage == FEEC_PC_STORE_CSR_HALF && s.P.nnz && !s.P16.p) kkt_refresh_half_csr(s);
  if (storage == FEEC_PC_STORE_SELL && (s.Psell.p || s.Psell_v16.p))
    bind_preconditioner(s, w, s.P32.p, nullptr, s.sdinv32.p, true);
  else if (storage != FEEC_PC_STORE_CSR && s.P16.p)
    bind_preconditioner(s, w, s.P32.p, s.P16.p, s.sdinv32.p);
  else
    bind_preconditioner(s, w, s.P32.p);
}

// FP64 solve with the polynomial in FP32 on the FP16 SELL copy (the mixed
// solve's preconditioner without its restarts). Without a polynomial, or
// without the SELL copy, the FP64 binding on P's CSR values.
static void bind_pc32(KktSystem& s, MinresWork<double>& w) {
  if (!s.val32.p) kkt_refresh_fp32(s);
  if (s.hx) {
    bind_hx(s, w, true);
    return;
  }
  w.hx = nullptr;
  const int m = (s.P.nnz > 0 && s.cheb_m > 0 && s.cheb.p) ? s.cheb_m : 0;
  if (m == 0 || !(s.Psell.p || s.Psell_v16.p) || !s.sdinv32.p) {
    bind_preconditioner(s, w, s.P.val.p);
    return;
  }
  SellF16 sel{};
  SellF16W selw{};
  if (s.sell_wide)
    selw = SellF16W{s.Psell_v16.p, s.Psell_col.p};
  else
    sel = SellF16{s.Psell.p};
  const int64_t n = std::max<int64_t>(w.n, 1);
  const bool realloc = (int64_t)w.res32.n != n;
  if (w.graph && (realloc || !w.pc32 || w.coef != s.cheb.p || w.cheb_m != m ||
                  w.sd32 != s.sdinv32.p || w.pval != s.P.val.p ||
                  std::memcmp(&w.sell16, &sel, sizeof(sel)) != 0 ||
                  std::memcmp(&w.sell16w, &selw, sizeof(selw)) != 0 ||
                  w.psell_ptr != s.Psell_ptr.p || w.sell_perm != s.sell_perm.p ||
                  w.sell_inv != s.sell_inv.p)) {
    (void)hipGraphExecDestroy(w.graph);
    w.graph = nullptr;
  }
  if (realloc) {
    w.res32.alloc(s.ctx, n);
    w.d0_32.alloc(s.ctx, n);
    w.d1_32.alloc(s.ctx, n);
  }
  w.pc32 = true;
  w.cheb_m = m;
  w.split = s.split;
  w.prp = s.P.row_ptr.p;
  w.pcol = s.P.col.p;
  w.pcol16 = s.Pcol16.p;
  w.pval = s.P.val.p;
  w.pval16 = nullptr;
  w.sd = nullptr;
  w.has_sell = false;
  w.sell_wide = s.sell_wide;
  w.sell = {};
  w.sellw = {};
  w.sell16 = sel;
  w.sell16w = selw;
  w.sd32 = s.sdinv32.p;
  w.a_nnz = s.A.nnz;
  w.p_entries = s.sell_entries;
  w.psell_ptr = s.Psell_ptr.p;
  w.sell_perm = s.sell_perm.p;
  w.sell_inv = s.sell_inv.p;
  w.coef = s.cheb.p;
}

int kkt_preconditioner_format(const KktSystem& s, int precision, int storage) {
  if (s.hx) return precision == FEEC_PREC_FP64 ? FEEC_PC_MULTIGRID_F64 : FEEC_PC_MULTIGRID_F32;
  if (precision == FEEC_PREC_FP64_PC32 && s.Psell_ptr.p && s.cheb_m > 0)
    return s.sell_wide ? FEEC_PC_SELL64_F16_C32 : FEEC_PC_SELL64_F16;
  const bool sell = storage == FEEC_PC_STORE_SELL && s.Psell_ptr.p;
  if (precision != FEEC_PREC_MIXED)
    return !sell ? FEEC_PC_CSR_F64 : s.sell_wide ? FEEC_PC_SELL64_F64_C32 : FEEC_PC_SELL64_F64;
  if (sell) return s.sell_wide ? FEEC_PC_SELL64_F16_C32 : FEEC_PC_SELL64_F16;
  if (storage != FEEC_PC_STORE_CSR && s.P.nnz) return FEEC_PC_CSR_F16;
  return FEEC_PC_CSR_F32;
}

// One solve as a sequence of MINRES runs with residual checks in between, so
// that a batch can run the MINRES part of many solves in lockstep. begin()
// sets up, prepare() fills w.b and the run's tolerance and budget (false when
// the solve is over), post() takes the run's result and checks the true
// residual, end() fills the report.
// MINRES monitors the residual in the preconditioner's norm. The first run of
// an FP64_PC32 solve stops at a quarter of rtol, so that the true residual
// usually meets rtol without a correction run: a restart costs more iterations
// than the few extra ones (3D Kuhn 64^3 k=1: 157 iterations in one run
// against 170 with a correction).
constexpr double kPc32FirstRun = 0.25;

struct SolveJob {
  KktSystem* s = nullptr;
  const double* b = nullptr;
  double* x = nullptr;
  feec_solve_opts o{};
  SolverState* ss = nullptr;
  feec_report rep{};
  double bnorm = 0.0, rnorm = 0.0, res = 1.0;
  int steps = 0;
  bool done = false;
  double tol = 0.0;
  int64_t budget = 0;
  // FP64_PC32: the next run continues the last one (resume_tol), once per solve
  bool resume = false, resumed = false;
  double resume_tol = 0.0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // FP64 Krylov (FEEC_PREC_FP64, FEEC_PREC_FP64_PC32) or FP32 refinement steps
  bool fp64() const { return o.precision != FEEC_PREC_MIXED; }
};

static void job_begin(SolveJob& j, SolverState** wsp) {
  KktSystem& s = *j.s;
  Ctx* ctx = s.ctx;
  hipStream_t st = ctx->stream;
  if (!*wsp) {
    *wsp = new SolverState();
    (*wsp)->host_s = static_cast<MinresS*>(ctx->pinned);
  }
  j.ss = *wsp;
  SolverState& ss = *j.ss;
  const int64_t n = s.n, nh = s.nh, N = n + nh;
  if ((int64_t)ss.r.n != N) {
    ss.r.alloc(ctx, N);
    ss.x.alloc(ctx, N);
    ss.part.alloc(ctx, NBLK * 8);
  }
  FEEC_HIP(hipEventCreate(&j.e0));
  FEEC_HIP(hipEventCreate(&j.e1));
  FEEC_HIP(hipEventRecord(j.e0, st));
  j.bnorm = device_norm(ctx, ss, j.b, N);
  FEEC_HIP(hipMemsetAsync(j.x, 0, N * sizeof(double), st));
  if (j.bnorm == 0.0) {
    j.done = true;
    j.rep.converged = 1;
    j.rep.residual = 0.0;
    j.rep.precond_residual = 0.0;
    j.res = 0.0;
    return;
  }
  if (j.fp64()) {
    MinresWork<double>& w = ss.w64;
    minres_alloc(ctx, w, n, nh);
    if (j.o.precision == FEEC_PREC_FP64_PC32)
      bind_pc32(s, w);
    else
      bind_fp64(s, w, j.o.precond_storage);
    FEEC_HIP(hipMemcpyAsync(w.dinv.p, s.inv_diag.p, N * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (nh) FEEC_HIP(hipMemcpyAsync(w.border.p, s.border.p, nh * n * sizeof(double), hipMemcpyDeviceToDevice, st));
    FEEC_HIP(hipMemcpyAsync(w.b.p, j.b, N * sizeof(double), hipMemcpyDeviceToDevice, st));
  } else {
    MinresWork<float>& w = ss.w32;
    minres_alloc(ctx, w, n, nh);
    bind_fp32(s, w, j.o.precond_storage);
    FEEC_HIP(hipMemcpyAsync(w.dinv.p, s.inv_diag32.p, N * sizeof(float), hipMemcpyDeviceToDevice, st));
    if (nh) k_f64_to_f32<<<grid_for(nh * n), TPB, 0, st>>>(s.border.p, nh * n, w.border.p);
    // r = b
    FEEC_HIP(hipMemcpyAsync(ss.r.p, j.b, N * sizeof(double), hipMemcpyDeviceToDevice, st));
    j.rnorm = j.bnorm;
  }
}

static bool job_prepare(SolveJob& j) {
  if (j.done) return false;
  const feec_solve_opts& o = j.o;
  if (!j.fp64() && j.steps >= o.max_refinements) {
    j.done = true;
    return false;
  }
  j.budget = o.max_iters - j.rep.iters;
  if (j.budget <= 0) {
    j.done = true;
    return false;
  }
  if (j.fp64()) {
    // FP64 MINRES on the reference's stopping quantity (FP64_PC32: its first
    // run to kPc32FirstRun rtol), then FP64 residual correction steps until
    // the true residual meets rtol. A correction only
    // has to close the remaining gap (relative to its own right-hand side,
    // whose norm is res * |b|), with a margin of 10 for the gap between the
    // preconditioner norm MINRES monitors and the 2-norm.
    const double first = o.precision == FEEC_PREC_FP64_PC32 ? kPc32FirstRun * o.rtol : o.rtol;
    j.tol = j.resume ? j.resume_tol : j.steps == 0 ? first : std::min(0.5, 0.1 * o.rtol / j.res);
    // a continued run gets at most half its previous length (plus 20) more
    if (j.resume) j.budget = std::min<int64_t>(j.budget, j.rep.iters / 2 + 20);
  } else {
    const int64_t N = j.s->n + j.s->nh;
    k_to_f32_scaled<<<grid_for(N), TPB, 0, j.s->ctx->stream>>>(j.ss->r.p, N, 1.0 / j.rnorm,
                                                                j.ss->w32.b.p);
    // the last refinement step only has to close the remaining gap (margin
    // of 10, as for the FP64 corrections)
    j.tol = std::max(o.inner_rtol, std::min(0.5, 0.1 * o.rtol / j.res));
  }
  return true;
}

static void job_post(SolveJob& j, const KrylovStats& ks) {
  KktSystem& s = *j.s;
  SolverState& ss = *j.ss;
  hipStream_t st = s.ctx->stream;
  const int64_t N = s.n + s.nh;
  j.rep.iters += ks.iters;
  j.rep.precond_residual = ks.precond_residual;
  if (j.fp64()) {
    k_axpy<<<grid_for(N), TPB, 0, st>>>(ss.w64.x.p, N, 1.0, j.x);
    j.res = true_residual(s, ss, j.b, j.x, j.bnorm);
    if (j.res <= j.o.rtol || j.steps >= j.o.max_refinements) {
      j.done = true;
      return;
    }
    ++j.steps;
    if (j.o.precision == FEEC_PREC_FP64_PC32 && !j.resumed && ks.converged && ks.iters > 0 &&
        ks.precond_residual > 0.0) {
      // The run met its target in the preconditioner norm but the true
      // residual is short by res / rtol: continue the same run to a target
      // lowered by that factor (margin 2) instead of restarting on the
      // residual, which would discard its Krylov space. Its iterate replaces
      // this one in x.
      k_axpy<<<grid_for(N), TPB, 0, st>>>(ss.w64.x.p, N, -1.0, j.x);
      j.resume = j.resumed = true;
      j.resume_tol = ks.precond_residual * std::min(0.5, 0.5 * j.o.rtol / j.res);
      return;
    }
    FEEC_HIP(hipMemcpyAsync(ss.w64.b.p, ss.r.p, N * sizeof(double), hipMemcpyDeviceToDevice, st));
  } else {
    k_axpy_from_f32<<<grid_for(N), TPB, 0, st>>>(ss.w32.x.p, N, j.rnorm, j.x);
    ++j.steps;
    j.res = true_residual(s, ss, j.b, j.x, j.bnorm);
    if (j.res <= j.o.rtol) {
      j.done = true;
      return;
    }
    j.rnorm = j.res * j.bnorm;
  }
}

static feec_report job_end(SolveJob& j) {
  if (j.bnorm != 0.0) {
    j.rep.refinements = j.steps;
    j.rep.residual = j.res;
    j.rep.converged = j.res <= j.o.rtol;
  }
  hipStream_t st = j.s->ctx->stream;
  FEEC_HIP(hipEventRecord(j.e1, st));
  FEEC_HIP(hipEventSynchronize(j.e1));
  float ms = 0.f;
  FEEC_HIP(hipEventElapsedTime(&ms, j.e0, j.e1));
  j.rep.seconds = ms * 1e-3;
  (void)hipEventDestroy(j.e0);
  (void)hipEventDestroy(j.e1);
  j.e0 = j.e1 = nullptr;
  return j.rep;
}

static KrylovStats job_run(SolveJob& j) {
  KktSystem& s = *j.s;
  SolverState& ss = *j.ss;
  if (j.fp64()) {
    const bool resume = j.resume;
    j.resume = false;
    return minres_run<double>(s.ctx, ss, ss.w64, s.A.row_ptr.p, s.A.col.p, s.A.val.p, j.tol,
                              j.budget, j.o, resume);
  }
  return minres_run<float>(s.ctx, ss, ss.w32, s.A.row_ptr.p, s.A.col.p, s.val32.p, j.tol, j.budget,
                           j.o);
}

// ---------------------------------------------------------------------------
// lockstep batch: the MINRES runs of many solves advance together, one launch
// per phase covering every active system. Each system gets exactly the blocks
// a standalone launch would give it (its own grid, the same block -> row map
// and the same partial sums), so its iterates are bit-identical to a
// standalone solve's. Systems whose run has finished drop out of the launch
// tables between segments.
// ---------------------------------------------------------------------------
__host__ __device__ inline int64_t red_blocks(int64_t n) {
  const int64_t b = (n + TPB - 1) / TPB;
  return b < 1 ? 1 : (b > NBLK ? NBLK : b);
}
__host__ __device__ inline int64_t spmv_blocks(int64_t n) {
  const int64_t b = (n + TPB / LANES - 1) / (TPB / LANES);
  return b < 1 ? 1 : (b > NBLK ? NBLK : b);
}
__host__ __device__ inline int64_t for_blocks(int64_t n) {
  const int64_t b = (n + TPB - 1) / TPB;
  return b < 1 ? 1 : b;
}

// TP: precision of the polynomial (float for FP64 solves with the FP32
// polynomial, else T)
template <class T, class TP = T>
struct BatchSys {
  int64_t n, nh, N, split;
  const int64_t* rp;
  const int32_t* col;
  const T* val;
  T *r1, *r2, *y, *z, *w, *w2, *x, *b, *border;
  TP *res, *d0, *d1;
  const TP* sd;
  const int32_t* perm;
  const int32_t* inv;
  typename MinresWork<TP>::SellN sell;
  typename MinresWork<TP>::SellW sellw;
  int wide, cheb_m;
  const int64_t* psell_ptr;
  const double* coef;
  double *part, *part2;
  MinresS* S;
  double rtol;
  long long max_iters;
  int par0;  // parity of the executed iterations when the run starts (a continued run may be odd)
};

// The system whose share of the launch holds this block (block starts are a
// prefix sum over the active systems; empty shares are skipped).
__device__ __forceinline__ int batch_find(const int64_t* starts, int ns) {
  const int64_t bid = blockIdx.x;
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (starts[mid] <= bid) lo = mid; else hi = mid - 1;
  }
  return lo;
}

#define FEEC_BATCH_SYS                                                                 \
  const int si_ = batch_find(starts, ns);                                              \
  const BatchSys<T, TP>& d = sys[si_];                                                     \
  const Blk bk{(int64_t)blockIdx.x - starts[si_], starts[si_ + 1] - starts[si_]};      \
  (void)bk;

template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) kb_zero(const BatchSys<T, TP>* sys, const int64_t* starts, int ns) {
  FEEC_BATCH_SYS
  k_mr_zero_body<T>(d.b, d.N, d.r1, d.r2, d.x, d.w, d.w2, bk);
}

// first polynomial application of a run (r = b into y; no MINRES scalars yet)
template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) kb_cheb3_init0(const BatchSys<T, TP>* sys, const int64_t* starts,
                                                      int ns) {
  FEEC_BATCH_SYS
  k_cheb3_init_body<T, TP>(d.b, d.sd, d.perm, d.res, d.d0, d.d1, d.n, d.split, d.coef,
                       (const MinresS*)nullptr, bk);
}

// step k of the polynomial; systems of lower degree sit it out
template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) kb_cheb3_step(const BatchSys<T, TP>* sys, const int64_t* starts,
                                                     int ns, int k, int with_s) {
  FEEC_BATCH_SYS
  if (k + 1 >= d.cheb_m) return;
  const TP* a = (k & 1) ? d.d0 : d.d1;
  TP* b = (k & 1) ? d.d1 : d.d0;
  const MinresS* S = with_s ? d.S : nullptr;
  if (d.wide)
    k_cheb3_step_body<TP, typename MinresWork<TP>::SellW>(d.sellw, d.psell_ptr, d.n, d.split, d.res,
                                                          a, b, d.coef, k, S, bk);
  else
    k_cheb3_step_body<TP, typename MinresWork<TP>::SellN>(d.sell, d.psell_ptr, d.n, d.split, d.res,
                                                          a, b, d.coef, k, S, bk);
}

// finish: initial (r = b, z = y, partials into part) or in-iteration (par)
template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) kb_cheb3_finish(const BatchSys<T, TP>* sys, const int64_t* starts,
                                                       int ns, int initial, int par) {
  FEEC_BATCH_SYS
  par ^= d.par0;
  const TP* x = ((d.cheb_m - 1) & 1) ? d.d0 : d.d1;
  if (initial)
    k_cheb3_finish_body<T, TP>(d.b, x, d.sd, d.inv, d.y, d.n, d.N, (const MinresS*)nullptr, d.part,
                               bk);
  else
    k_cheb3_finish_body<T, TP>(par ? d.r2 : d.r1, x, d.sd, d.inv, par ? d.y : d.z, d.n, d.N, d.S,
                               d.part2, bk);
}

template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) kb_start(const BatchSys<T, TP>* sys, const int64_t* starts, int ns) {
  FEEC_BATCH_SYS
  k_mr_start_body(d.S, d.part, (int)red_blocks(d.N), d.rtol, d.max_iters, bk);
}

template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) kb_spmv(const BatchSys<T, TP>* sys, const int64_t* starts, int ns,
                                               int par) {
  FEEC_BATCH_SYS
  par ^= d.par0;
  k_mr_spmv_body<T>(d.rp, d.col, d.val, d.n, par ? d.z : d.y, par ? d.r2 : d.r1, d.S, d.part, bk);
}

template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) kb_border_dot(const BatchSys<T, TP>* sys, const int64_t* starts,
                                                     int ns, int par) {
  FEEC_BATCH_SYS
  par ^= d.par0;
  k_mr_border_dot_body<T>(d.border, (int)d.nh, d.n, par ? d.z : d.y, d.S, d.part + NBLK, bk);
}

template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) kb_border_rows(const BatchSys<T, TP>* sys, const int64_t* starts,
                                                      int ns, int par) {
  FEEC_BATCH_SYS
  par ^= d.par0;
  k_mr_border_rows_body<T>((int)d.nh, d.n, par ? d.z : d.y, par ? d.r2 : d.r1, d.S, d.part + NBLK,
                           (int)red_blocks(d.n), bk);
}

template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) kb_border_axpy(const BatchSys<T, TP>* sys, const int64_t* starts,
                                                      int ns, int par) {
  FEEC_BATCH_SYS
  par ^= d.par0;
  k_mr_border_axpy_body<T>(d.border, (int)d.nh, d.n, par ? d.z : d.y, par ? d.r2 : d.r1, d.S, bk);
}

template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) kb_alpha(const BatchSys<T, TP>* sys, const int64_t* starts, int ns) {
  FEEC_BATCH_SYS
  k_mr_alpha_body(d.S, d.part, (int)spmv_blocks(d.n), d.nh ? 1 : 0, bk);
}

template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) kb_update_cheb3_init(const BatchSys<T, TP>* sys,
                                                            const int64_t* starts, int ns, int par) {
  FEEC_BATCH_SYS
  par ^= d.par0;
  k_mr_update_cheb3_init_body<T, TP>(par ? d.r2 : d.r1, par ? d.r1 : d.r2, d.sd, d.perm, d.res, d.d0,
                                 d.d1, d.n, d.N, d.split, d.coef, d.S, bk);
}

template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) kb_rotate(const BatchSys<T, TP>* sys, const int64_t* starts, int ns) {
  FEEC_BATCH_SYS
  k_mr_rotate_body(d.S, d.part2, (int)red_blocks(d.N), bk);
}

template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) kb_wx(const BatchSys<T, TP>* sys, const int64_t* starts, int ns,
                                             int par) {
  FEEC_BATCH_SYS
  par ^= d.par0;
  k_mr_wx_body<T>(par ? d.z : d.y, par ? d.w2 : d.w, par ? d.w : d.w2, d.x, d.N, d.S, bk);
}

template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) kb_check(const BatchSys<T, TP>* sys, const int64_t* starts, int ns) {
  FEEC_BATCH_SYS
  if (threadIdx.x == 0) k_mr_check_body(d.S, bk);
}

// continue each system's finished run to its new target (k_mr_resume)
template <class T, class TP = T>
__global__ void __launch_bounds__(TPB) kb_resume(const BatchSys<T, TP>* sys, const int64_t* starts,
                                                 int ns) {
  FEEC_BATCH_SYS
  if (threadIdx.x == 0) {
    d.S->done = 0;
    d.S->rtol = d.rtol;
    d.S->max_iters = d.max_iters;
  }
}

template <class T, class TP = T>
__global__ void kb_gather_s(const BatchSys<T, TP>* sys, int ns, MinresS* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ns) out[i] = *sys[i].S;
}

template <class T, class TP = T>
static BatchSys<T, TP> batch_sys(MinresWork<T>& w, KktSystem& s, const T* val, double rtol,
                                 int64_t max_iters) {
  BatchSys<T, TP> d{};
  d.n = w.n;
  d.nh = w.nh;
  d.N = w.n + w.nh;
  d.split = w.split;
  d.rp = s.A.row_ptr.p;
  d.col = s.A.col.p;
  d.val = val;
  d.r1 = w.r1.p; d.r2 = w.r2.p; d.y = w.y.p; d.z = w.z.p; d.w = w.w.p; d.w2 = w.w2.p;
  d.x = w.x.p; d.b = w.b.p; d.border = w.border.p;
  if constexpr (std::is_same_v<T, TP>) {
    d.res = w.res.p; d.d0 = w.d0.p; d.d1 = w.d1.p;
    d.sd = w.sd;
    d.sell = w.sell;
    d.sellw = w.sellw;
  } else {  // FP64 solve, FP32 polynomial on the FP16 copy
    d.res = w.res32.p; d.d0 = w.d0_32.p; d.d1 = w.d1_32.p;
    d.sd = w.sd32;
    d.sell = w.sell16;
    d.sellw = w.sell16w;
  }
  d.perm = w.sell_perm;
  d.inv = w.sell_inv;
  d.wide = w.sell_wide ? 1 : 0;
  d.cheb_m = w.cheb_m;
  d.psell_ptr = w.psell_ptr;
  d.coef = w.coef;
  d.part = w.part.p;
  d.part2 = w.part2.p;
  d.S = w.S.p;
  d.rtol = rtol;
  d.max_iters = (long long)max_iters;
  return d;
}

// Launch tables of one set of systems: the BatchSys array and, per grid
// shape, the prefix sums of the systems' block counts.
template <class T, class TP = T>
struct BatchTables {
  enum { RED_N, RED_n, SPMV, FOR_n, FOR_N, ONE, BORDER_RED, BORDER_ONE, NSHAPE };
  int ns = 0;
  int max_m = 0;
  bool any_border = false;
  DBuf<BatchSys<T, TP>> sys;
  DBuf<int64_t> starts;  // [NSHAPE][ns + 1]
  int64_t total[NSHAPE] = {};
  std::vector<std::vector<int64_t>> hstart;  // host copy of the prefix sums per shape
  std::vector<int> ns_step;                 // systems (a prefix) whose degree reaches step k
  void build(Ctx* ctx, const std::vector<BatchSys<T, TP>>& v) {
    ns = (int)v.size();
    max_m = 0;
    any_border = false;
    std::vector<int64_t> h((size_t)NSHAPE * (ns + 1), 0);
    for (int i = 0; i < ns; ++i) {
      const BatchSys<T, TP>& d = v[i];
      max_m = std::max(max_m, d.cheb_m);
      any_border = any_border || d.nh > 0;
      const int64_t g[NSHAPE] = {red_blocks(d.N), red_blocks(d.n), spmv_blocks(d.n), for_blocks(d.n),
                                 for_blocks(d.N), 1, d.nh ? red_blocks(d.n) : 0, d.nh ? 1 : 0};
      for (int k = 0; k < NSHAPE; ++k) h[(size_t)k * (ns + 1) + i + 1] = h[(size_t)k * (ns + 1) + i] + g[k];
    }
    for (int k = 0; k < NSHAPE; ++k) total[k] = h[(size_t)k * (ns + 1) + ns];
    hstart.assign(NSHAPE, {});
    for (int k = 0; k < NSHAPE; ++k)
      hstart[k].assign(h.begin() + (size_t)k * (ns + 1), h.begin() + (size_t)(k + 1) * (ns + 1));
    ns_step.assign(std::max(max_m, 1), 0);
    for (int k = 0; k + 1 < max_m; ++k) {
      int c = 0;
      while (c < ns && v[c].cheb_m > k + 1) ++c;  // v is sorted by decreasing degree
      ns_step[k] = c;
    }
    sys.alloc(ctx, std::max(ns, 1));
    starts.alloc(ctx, h.size());
    copy_to_device(ctx, sys.p, v.data(), v.size() * sizeof(BatchSys<T, TP>));
    copy_to_device(ctx, starts.p, h.data(), h.size() * sizeof(int64_t));
  }
  const int64_t* st(int k) const { return starts.p + (size_t)k * (ns + 1); }
};

// Step k over the leading systems (sorted by decreasing degree) whose degree
// reaches it.
template <class T, class TP>
static void batch_steps(hipStream_t st, const BatchTables<T, TP>& t, int k, int with_s) {
  using Tab = BatchTables<T, TP>;
  const int ns_k = t.ns_step[k];
  if (ns_k <= 0) return;
  const int64_t blocks = t.hstart[Tab::FOR_n][ns_k];
  if (blocks <= 0) return;
  kb_cheb3_step<T, TP><<<(unsigned)blocks, TPB, 0, st>>>(t.sys.p, t.st(Tab::FOR_n), ns_k, k,
                                                         with_s);
}

template <class T, class TP, class K, class... A>
static void batch_launch(hipStream_t st, K kernel, const BatchTables<T, TP>& t, int shape, A... a) {
  const int64_t blocks = t.total[shape];
  if (blocks <= 0 || t.ns == 0) return;
  kernel<<<(unsigned)blocks, TPB, 0, st>>>(t.sys.p, t.st(shape), t.ns, a...);
}

// One MINRES run (w.b filled, preconditioner bound, SELL storage and a
// polynomial of degree >= 2) of each system, in lockstep on ctx's stream.
// With `resume`, each system's finished run continues instead (rtol: its new
// target; max_iters: its further iterations), with the buffer parity its
// executed iterations left.
template <class T, class TP = T>
static void minres_lockstep(Ctx* ctx, std::vector<BatchSys<T, TP>>& all,
                            std::vector<KrylovStats>& out, bool resume = false) {
  using Tab = BatchTables<T, TP>;
  hipStream_t st = ctx->stream;
  out.assign(all.size(), KrylovStats{});
  if (all.empty()) return;
  std::vector<long long> start_iter(all.size(), 0);
  if (resume) {
    for (size_t i = 0; i < all.size(); ++i) {
      MinresS h{};
      copy_to_host(ctx, &h, all[i].S, sizeof(h));
      start_iter[i] = h.iter;
      all[i].par0 = (int)(h.iter & 1);
      all[i].max_iters += h.iter;
    }
  }
  // systems in order of decreasing degree: step k of the polynomial then
  // launches only over the leading systems whose degree reaches it
  std::vector<int> order(all.size());
  for (size_t i = 0; i < all.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return all[a].cheb_m > all[b].cheb_m; });
  std::vector<BatchSys<T, TP>> sorted;
  for (int i : order) sorted.push_back(all[i]);
  Tab t;
  t.build(ctx, sorted);
  if (resume) {
    batch_launch(st, kb_resume<T, TP>, t, Tab::ONE);
  } else {
    // start of the run: r1 = r2 = b, y = B b, beta1
    batch_launch(st, kb_zero<T, TP>, t, Tab::RED_N);
    batch_launch(st, kb_cheb3_init0<T, TP>, t, Tab::RED_n);
    for (int k = 0; k + 1 < t.max_m; ++k) batch_steps(st, t, k, 0);
    batch_launch(st, kb_cheb3_finish<T, TP>, t, Tab::RED_N, 1, 0);
    batch_launch(st, kb_start<T, TP>, t, Tab::ONE);
  }
  FEEC_HIP(hipGetLastError());
  std::vector<int> active = order;
  DBuf<MinresS> hs_dev(ctx, all.size());
  std::vector<MinresS> hs(all.size());
  constexpr int SEG = 16;
  while (!active.empty()) {
    for (int it = 0; it < SEG; ++it) {
      const int par = it & 1;
      batch_launch(st, kb_spmv<T, TP>, t, Tab::SPMV, par);
      if (t.any_border) {
        batch_launch(st, kb_border_dot<T, TP>, t, Tab::BORDER_RED, par);
        batch_launch(st, kb_border_rows<T, TP>, t, Tab::BORDER_ONE, par);
        batch_launch(st, kb_border_axpy<T, TP>, t, Tab::BORDER_RED, par);
      }
      batch_launch(st, kb_alpha<T, TP>, t, Tab::ONE);
      batch_launch(st, kb_update_cheb3_init<T, TP>, t, Tab::RED_N, par);
      for (int k = 0; k + 1 < t.max_m; ++k) batch_steps(st, t, k, 1);
      batch_launch(st, kb_cheb3_finish<T, TP>, t, Tab::RED_N, 0, par);
      batch_launch(st, kb_rotate<T, TP>, t, Tab::ONE);
      batch_launch(st, kb_wx<T, TP>, t, Tab::FOR_N, par);
      batch_launch(st, kb_check<T, TP>, t, Tab::ONE);
    }
    FEEC_HIP(hipGetLastError());
    kb_gather_s<T, TP><<<(unsigned)((t.ns + TPB - 1) / TPB), TPB, 0, st>>>(t.sys.p, t.ns, hs_dev.p);
    FEEC_HIP(hipGetLastError());
    copy_to_host(ctx, hs.data(), hs_dev.p, t.ns * sizeof(MinresS));
    std::vector<int> still;
    std::vector<BatchSys<T, TP>> v;
    for (int i = 0; i < t.ns; ++i) {
      const int id = active[i];
      if (hs[i].done) {
        out[id].iters = hs[i].iter - start_iter[id];
        out[id].precond_residual = hs[i].residual;
        out[id].converged = hs[i].residual <= all[id].rtol;
      } else {
        still.push_back(id);
        v.push_back(all[id]);
      }
    }
    if (still.size() != active.size() && !still.empty()) t.build(ctx, v);
    active.swap(still);
  }
}

// Whether a solve's MINRES runs can go through the lockstep path.
static bool lockstep_eligible(const SolveJob& j) {
  if (!j.ss) return false;
  if (j.fp64()) {
    const MinresWork<double>& w = j.ss->w64;
    return ((w.has_sell && w.sd) || w.pc32) && w.cheb_m >= 2;
  }
  return j.ss->w32.has_sell && j.ss->w32.sd && j.ss->w32.cheb_m >= 2;
}

void kkt_solve_batch(std::vector<KktSystem*>& systems, std::vector<const double*>& b,
                     std::vector<double*>& x, const std::vector<feec_solve_opts>& opts,
                     std::vector<SolverState**>& ws, std::vector<feec_report>& reps, Ctx* lead) {
  const size_t m = systems.size();
  std::vector<SolveJob> jobs(m);
  for (size_t i = 0; i < m; ++i) {
    jobs[i].s = systems[i];
    jobs[i].b = b[i];
    jobs[i].x = x[i];
    jobs[i].o = opts[i];
    job_begin(jobs[i], ws[i]);
  }
  while (true) {
    std::vector<size_t> run32, run64, runpc, runpc_resume, solo;
    for (size_t i = 0; i < m; ++i) {
      if (!job_prepare(jobs[i])) continue;
      if (!lockstep_eligible(jobs[i]))
        solo.push_back(i);
      else if (jobs[i].fp64())
        (jobs[i].ss->w64.pc32 ? (jobs[i].resume ? runpc_resume : runpc) : run64).push_back(i);
      else
        run32.push_back(i);
    }
    if (run32.empty() && run64.empty() && runpc.empty() && runpc_resume.empty() && solo.empty())
      break;
    for (size_t i : solo) job_post(jobs[i], job_run(jobs[i]));
    // the systems' own streams have filled w.b; the lockstep runs on lead's
    for (size_t i : run32) jobs[i].s->ctx->sync();
    for (size_t i : run64) jobs[i].s->ctx->sync();
    for (size_t i : runpc) jobs[i].s->ctx->sync();
    for (size_t i : runpc_resume) jobs[i].s->ctx->sync();
    std::vector<KrylovStats> ks;
    if (!run32.empty()) {
      std::vector<BatchSys<float>> v;
      for (size_t i : run32)
        v.push_back(batch_sys<float>(jobs[i].ss->w32, *jobs[i].s, jobs[i].s->val32.p, jobs[i].tol,
                                     jobs[i].budget));
      minres_lockstep<float>(lead, v, ks);
      lead->sync();
      for (size_t a = 0; a < run32.size(); ++a) job_post(jobs[run32[a]], ks[a]);
    }
    if (!run64.empty()) {
      std::vector<BatchSys<double>> v;
      for (size_t i : run64)
        v.push_back(batch_sys<double>(jobs[i].ss->w64, *jobs[i].s, jobs[i].s->A.val.p, jobs[i].tol,
                                      jobs[i].budget));
      minres_lockstep<double>(lead, v, ks);
      lead->sync();
      for (size_t a = 0; a < run64.size(); ++a) job_post(jobs[run64[a]], ks[a]);
    }
    if (!runpc.empty()) {
      std::vector<BatchSys<double, float>> v;
      for (size_t i : runpc)
        v.push_back(batch_sys<double, float>(jobs[i].ss->w64, *jobs[i].s, jobs[i].s->A.val.p,
                                             jobs[i].tol, jobs[i].budget));
      minres_lockstep<double, float>(lead, v, ks);
      lead->sync();
      for (size_t a = 0; a < runpc.size(); ++a) job_post(jobs[runpc[a]], ks[a]);
    }
    if (!runpc_resume.empty()) {
      std::vector<BatchSys<double, float>> v;
      for (size_t i : runpc_resume) {
        jobs[i].resume = false;
        v.push_back(batch_sys<double, float>(jobs[i].ss->w64, *jobs[i].s, jobs[i].s->A.val.p,
                                             jobs[i].tol, jobs[i].budget));
      }
      minres_lockstep<double, float>(lead, v, ks, true);
      lead->sync();
      for (size_t a = 0; a < runpc_resume.size(); ++a) job_post(jobs[runpc_resume[a]], ks[a]);
    }
  }
  reps.resize(m);
  for (size_t i = 0; i < m; ++i) reps[i] = job_end(jobs[i]);
}

feec_report kkt_solve(KktSystem& s, const double* b, double* x_out, const feec_solve_opts& o,
                      SolverState** wsp) {
  SolveJob j;
  j.s = &s;
  j.b = b;
  j.x = x_out;
  j.o = o;
  job_begin(j, wsp);
  while (job_prepare(j)) job_post(j, job_run(j));
  return job_end(j);
}

// Mean seconds per launch of `reps` launches replayed from one graph, as the
// solve replays its kernels (after one warm-up replay).
template <class Build>
static double time_graph(hipStream_t st, int reps, Build&& build) {
  reps = std::max(reps, 1);
  hipGraph_t g = nullptr;
  FEEC_HIP(hipGraphCreate(&g, 0));
  std::vector<std::shared_ptr<void>> keep;
  hipGraphExec_t ex = nullptr;
  try {
    Enq gq{st, g, &keep};
    for (int i = 0; i < reps; ++i) build(gq);
    FEEC_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
  } catch (...) {
    (void)hipGraphDestroy(g);
    throw;
  }
  (void)hipGraphDestroy(g);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.f;
  hipError_t e = hipGraphLaunch(ex, st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  if (e == hipSuccess) e = hipGraphLaunch(ex, st);
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipGraphExecDestroy(ex);
  FEEC_HIP(e);
  return ms * 1e-3 / reps;
}

// Mean duration of the fused MINRES SpMV launch (the kernel that moves the
// matrix) on this system, FP32 or FP64, replayed from a graph as in the solve.
double time_minres_spmv(KktSystem& s, bool fp32, int reps) {
  Ctx* ctx = s.ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = s.n, N = s.n + s.nh;
  if (fp32 && !s.val32.p) kkt_refresh_fp32(s);
  DBuf<MinresS> S(ctx, 1);
  MinresS h{};
  h.beta = 1.0;
  h.oldb = 1.0;
  h.iter = 1;
  h.done = 0;
  copy_to_device(ctx, S.p, &h, sizeof(h));
  DBuf<double> part(ctx, NBLK);
  if (fp32) {
    DBuf<float> y(ctx, N), r(ctx, N);
    FEEC_HIP(hipMemsetAsync(y.p, 0, N * 4, st));
    FEEC_HIP(hipMemsetAsync(r.p, 0, N * 4, st));
    return time_graph(st, reps, [&](Enq& q) {
      q(k_mr_spmv<float>, spmv_grid(n), TPB, s.A.row_ptr.p, s.A.col.p, s.val32.p, n, y.p, r.p, S.p,
        part.p);
    });
  }
  DBuf<double> y(ctx, N), r(ctx, N);
  FEEC_HIP(hipMemsetAsync(y.p, 0, N * 8, st));
  FEEC_HIP(hipMemsetAsync(r.p, 0, N * 8, st));
  return time_graph(st, reps, [&](Enq& q) {
    q(k_mr_spmv<double>, spmv_grid(n), TPB, s.A.row_ptr.p, s.A.col.p, s.A.val.p, n, y.p, r.p, S.p,
      part.p);
  });
}

// z = B r, B the block preconditioner exactly as a solve of the given
// precision with `storage` (feec_pc_storage) applies it; r and z are N device
// doubles.
void kkt_apply_preconditioner(KktSystem& s, int precision, int storage, const double* r,
                              double* z) {
  Ctx* ctx = s.ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = s.n, N = s.n + s.nh;
  if (precision != FEEC_PREC_MIXED) {
    MinresWork<double> w;
    minres_alloc(ctx, w, n, s.nh);
    if (precision == FEEC_PREC_FP64_PC32)
      bind_pc32(s, w);
    else
      bind_fp64(s, w, storage);
    FEEC_HIP(hipMemcpyAsync(w.dinv.p, s.inv_diag.p, N * sizeof(double), hipMemcpyDeviceToDevice, st));
    FEEC_HIP(hipMemcpyAsync(w.b.p, r, N * sizeof(double), hipMemcpyDeviceToDevice, st));
    Enq q{st};
    if (w.hx)
      hx_enqueue(q, *w.hx, (const double*)w.b.p, w.y.p, nullptr, w.part.p, w.hx_fp32);
    else if (w.cheb_m > 0)
      cheb_apply<double>(q, w, w.b.p, w.y.p, nullptr, w.part.p);
    else
      k_jacobi<<<grid_for(N), TPB, 0, st>>>(w.b.p, w.dinv.p, w.y.p, N);
    FEEC_HIP(hipMemcpyAsync(z, w.y.p, N * sizeof(double), hipMemcpyDeviceToDevice, st));
    ctx->sync();
    return;
  }
  MinresWork<float> w;
  minres_alloc(ctx, w, n, s.nh);
  bind_fp32(s, w, storage);
  FEEC_HIP(hipMemcpyAsync(w.dinv.p, s.inv_diag32.p, N * sizeof(float), hipMemcpyDeviceToDevice, st));
  k_f64_to_f32<<<grid_for(N), TPB, 0, st>>>(r, N, w.b.p);
  Enq q{st};
  if (w.hx)
    hx_enqueue(q, *w.hx, (const float*)w.b.p, w.y.p, nullptr, w.part.p);
  else if (w.cheb_m > 0)
    cheb_apply<float>(q, w, w.b.p, w.y.p, nullptr, w.part.p);
  else
    k_jacobi32<<<grid_for(N), TPB, 0, st>>>(w.b.p, w.dinv.p, w.y.p, N);
  FEEC_HIP(hipMemsetAsync(z, 0, N * sizeof(double), st));
  k_axpy_from_f32<<<grid_for(N), TPB, 0, st>>>(w.y.p, N, 1.0, z);
  FEEC_HIP(hipGetLastError());
  ctx->sync();
}

// Mean duration of one Chebyshev step (the preconditioner's fused SpMV) on
// this system with the given storage, FP32 or FP64, replayed from a graph as
// in the solve; 0 when the system has no polynomial.
double time_cheb_step(KktSystem& s, bool fp32, int reps, int storage) {
  if (s.P.nnz == 0 || s.cheb_m < 2 || !s.cheb.p) return 0.0;
  Ctx* ctx = s.ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = s.n, N = s.n + s.nh;
  auto run = [&](auto& w) {
    using T = std::remove_reference_t<decltype(*w.res.p)>;
    for (T* v : {w.res.p, w.d0.p, w.d1.p, w.z.p})
      FEEC_HIP(hipMemsetAsync(v, 0, N * sizeof(T), st));
    const MinresS* none = nullptr;
    return time_graph(st, reps, [&](Enq& q) { cheb_step<T>(q, w, w.d0.p, w.d1.p, w.z.p, 0, none); });
  };
  if (fp32) {
    MinresWork<float> w;
    minres_alloc(ctx, w, n, s.nh);
    bind_fp32(s, w, storage);
    return run(w);
  }
  MinresWork<double> w;
  minres_alloc(ctx, w, n, s.nh);
  bind_fp64(s, w, storage);
  return run(w);
}

void spmv64(Ctx* ctx, const Csr& a, const double* x, double* y) {
  if (a.nrows == 0) return;
  k_spmv<double><<<spmv_grid(a.nrows), TPB, 0, ctx->stream>>>(a.row_ptr.p, a.col.p, a.val.p,
                                                              a.nrows, x, y);
  FEEC_HIP(hipGetLastError());
}

// a . b over n device doubles (fixed-order partial sums), read on the host.
namespace {
__global__ void k_spmv_rowseq(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                              const double* __restrict__ val, int64_t n, const double* __restrict__ x,
                              double* __restrict__ y) {
#pragma clang fp contract(off)
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  double acc = 0.0;
  for (int64_t p = rp[r]; p < rp[r + 1]; ++p) acc = acc + val[p] * x[col[p]];
  y[r] = acc;
}
}  // namespace

void spmv_rowseq(Ctx* ctx, const Csr& a, const double* x, double* y) {
  if (a.nrows == 0) return;
  k_spmv_rowseq<<<grid_for(a.nrows), TPB, 0, ctx->stream>>>(a.row_ptr.p, a.col.p, a.val.p, a.nrows,
                                                            x, y);
  FEEC_HIP(hipGetLastError());
}

double dot64(Ctx* ctx, const double* a, const double* b, int64_t n, double* part) {
  hipStream_t st = ctx->stream;
  k_dot<double><<<red_grid(n), TPB, 0, st>>>(a, b, n, part);
  k_finish_sum<<<1, TPB, 0, st>>>(part, red_grid(n), part + NBLK);
  double h = 0.0;
  FEEC_HIP(hipMemcpyAsync(&h, part + NBLK, sizeof(double), hipMemcpyDeviceToHost, st));
  FEEC_HIP(hipStreamSynchronize(st));
  return h;
}

// x += s a on n device doubles.
void axpy64(Ctx* ctx, const double* a, int64_t n, double s, double* x) {
  if (n > 0) k_axpy<<<grid_for(n), TPB, 0, ctx->stream>>>(a, n, s, x);
  FEEC_HIP(hipGetLastError());
}

// --- device-resident Jacobi CG -------------------------------------------------
// The scalars live in device memory; the host reads them once per segment of
// CG_SEG iterations, and iterations after convergence return at once.
struct CgS {
  double rz, bnorm2, rtol2, alpha, beta, res2;
  long long iter, max_iters;
  int done;
  int pad;
};
constexpr int CG_SEG = 8;

// x = 0, r = b, z = p = dinv b; partials r.z and b.b
__global__ void __launch_bounds__(TPB) k_cg_init(const double* __restrict__ b,
                                                 const double* __restrict__ dinv, int64_t n,
                                                 double* __restrict__ x, double* __restrict__ r,
                                                 double* __restrict__ z, double* __restrict__ p,
                                                 double* part) {
  __shared__ double sh[TPB / 64];
  double rz = 0.0, bb = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double bi = b[i], zi = dinv[i] * bi;
    x[i] = 0.0;
    r[i] = bi;
    z[i] = zi;
    p[i] = zi;
    rz += bi * zi;
    bb += bi * bi;
  }
  const double t0 = block_reduce(rz, sh);
  __syncthreads();
  const double t1 = block_reduce(bb, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = t0;
    part[NBLK + blockIdx.x] = t1;
  }
}

__global__ void k_cg_start(CgS* S, const double* part, int np, double rtol, long long max_iters) {
  __shared__ double sh[TPB / 64];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    a += part[i];
    b += part[NBLK + i];
  }
  const double rz = block_reduce(a, sh);
  __syncthreads();
  const double bb = block_reduce(b, sh);
  if (threadIdx.x == 0) {
    CgS s{};
    s.rz = rz;
    s.bnorm2 = bb;
    s.rtol2 = rtol * rtol;
    s.res2 = bb;
    s.max_iters = max_iters;
    s.done = (bb == 0.0 || max_iters <= 0) ? 1 : 0;
    *S = s;
  }
}

// ap = A p; partial p.ap
__global__ void __launch_bounds__(TPB) k_cg_spmv(const int64_t* __restrict__ rp,
                                                 const int32_t* __restrict__ col,
                                                 const double* __restrict__ val, int64_t n,
                                                 const double* __restrict__ p,
                                                 double* __restrict__ ap, const CgS* S,
                                                 double* part) {
  if (S->done) return;
  __shared__ double sh[TPB / 64];
  const int lane = threadIdx.x % LANES;
  const int64_t groups = (int64_t)gridDim.x * (TPB / LANES);
  double loc = 0.0;
  for (int64_t row = blockIdx.x * (int64_t)(TPB / LANES) + threadIdx.x / LANES; row < n;
       row += groups) {
    const double acc = row_dot<double>(rp, col, val, p, row, lane);
    if (lane == 0) {
      ap[row] = acc;
      loc += p[row] * acc;
    }
  }
  const double t = block_reduce(loc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void k_cg_alpha(CgS* S, const double* part, int np) {
  if (S->done) return;
  __shared__ double sh[TPB / 64];
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) v += part[i];
  const double pap = block_reduce(v, sh);
  if (threadIdx.x == 0) S->alpha = S->rz / pap;
}

// x += alpha p, r -= alpha ap, z = dinv r; partials r.z and r.r
__global__ void __launch_bounds__(TPB) k_cg_update(double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ z,
                                                   const double* __restrict__ p,
                                                   const double* __restrict__ ap,
                                                   const double* __restrict__ dinv, int64_t n,
                                                   const CgS* S, double* part) {
  if (S->done) return;
  __shared__ double sh[TPB / 64];
  const double alpha = S->alpha;
  double rz = 0.0, rr = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * ap[i], zi = dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
  }
  const double t0 = block_reduce(rz, sh);
  __syncthreads();
  const double t1 = block_reduce(rr, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = t0;
    part[NBLK + blockIdx.x] = t1;
  }
}

__global__ void k_cg_beta(CgS* S, const double* part, int np) {
  if (S->done) return;
  __shared__ double sh[TPB / 64];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    a += part[i];
    b += part[NBLK + i];
  }
  const double rz = block_reduce(a, sh);
  __syncthreads();
  const double rr = block_reduce(b, sh);
  if (threadIdx.x == 0) {
    CgS s = *S;
    s.beta = rz / s.rz;
    s.rz = rz;
    s.res2 = rr;
    s.iter += 1;
    if (rr <= s.rtol2 * s.bnorm2 || s.iter >= s.max_iters) s.done = 1;
    *S = s;
  }
}

// p = z + beta p
__global__ void k_cg_dir(double* __restrict__ p, const double* __restrict__ z, int64_t n,
                         const CgS* S) {
  if (S->done) return;
  const double beta = S->beta;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

// Jacobi CG (krylov.rs:48-95: stop when |r| <= rtol |b|) with its scalars on
// the device: five launches per iteration, one host read per CG_SEG iterations.
namespace {

// Device buffers of one CG run.
struct CgWork {
  double *r, *z, *p, *ap, *part;
  CgS* S;
};

KrylovStats cg_solve_ws(Ctx* ctx, const Csr& a, const double* dinv, const double* b, double* x,
                        double rtol, int64_t max_iters, const CgWork& w) {
  const int64_t n = a.nrows;
  hipStream_t st = ctx->stream;
  KrylovStats ks;
  if (n == 0) {
    ks.converged = true;
    ks.precond_residual = 0.0;
    return ks;
  }
  const unsigned gr = red_grid(n), gs = spmv_grid(n);
  k_cg_init<<<gr, TPB, 0, st>>>(b, dinv, n, x, w.r, w.z, w.p, w.part);
  k_cg_start<<<1, TPB, 0, st>>>(w.S, w.part, (int)gr, rtol, (long long)max_iters);
  CgS h{};
  while (true) {
    for (int it = 0; it < CG_SEG; ++it) {
      k_cg_spmv<<<gs, TPB, 0, st>>>(a.row_ptr.p, a.col.p, a.val.p, n, w.p, w.ap, w.S, w.part);
      k_cg_alpha<<<1, TPB, 0, st>>>(w.S, w.part, (int)gs);
      k_cg_update<<<gr, TPB, 0, st>>>(x, w.r, w.z, w.p, w.ap, dinv, n, w.S, w.part);
      k_cg_beta<<<1, TPB, 0, st>>>(w.S, w.part, (int)gr);
      k_cg_dir<<<gr, TPB, 0, st>>>(w.p, w.z, n, w.S);
    }
    FEEC_HIP(hipGetLastError());
    copy_to_host(ctx, &h, w.S, sizeof(h));
    if (h.done) break;
  }
  ks.iters = h.iter;
  ks.precond_residual = h.bnorm2 > 0.0 ? std::sqrt(h.res2 / h.bnorm2) : 0.0;
  ks.converged = h.bnorm2 == 0.0 || h.res2 <= h.rtol2 * h.bnorm2;
  return ks;
}

// ---------------------------------------------------------------------------
// feec_csr_cg_noalloc's scratch: a file-scope device array, one copy per GPU
// (each at its own address), held until the process ends. Layout, in doubles:
// the CG scalars (padded to 256 B), 2 * NBLK partial sums, then seven vectors
// of kNoAllocMaxRows: b, x (staging), the inverse diagonal, r, z, p, A p.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocMaxRows = int64_t(1) << 18;
constexpr int64_t kScalarDoubles = 32;
static_assert(sizeof(CgS) <= kScalarDoubles * sizeof(double), "CG scalars fit their slot");
constexpr int64_t kNoAllocDoubles = kScalarDoubles + 2 * NBLK + 7 * kNoAllocMaxRows;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_ptr;  // device -> that device's copy of `scratch`

double* scratch_for(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  return it == g_scratch_ptr.end() ? nullptr : it->second;
}

__global__ void k_reciprocal_or_one(const double* d, int64_t n, double* out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = d[i] != 0.0 ? 1.0 / d[i] : 1.0;
}

}  // namespace

void noalloc_scratch_init(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr.count(device)) return;
  void* p = nullptr;
  FEEC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  FEEC_REQUIRE(p, FEEC_E_DEVICE, "no device address for the scratch array");
  g_scratch_ptr[device] = static_cast<double*>(p);
}

KrylovStats cg_solve_device(Ctx* ctx, const Csr& a, const double* dinv, const double* b, double* x,
                            double rtol, int64_t max_iters) {
  const int64_t n = std::max<int64_t>(a.nrows, 1);
  DBuf<double> r(ctx, n), z(ctx, n), p(ctx, n), ap(ctx, n), part(ctx, 2 * NBLK);
  DBuf<CgS> S(ctx, 1);
  return cg_solve_ws(ctx, a, dinv, b, x, rtol, max_iters, CgWork{r.p, z.p, p.p, ap.p, part.p, S.p});
}

}  // namespace feec

using namespace feec;

extern "C" {

// The Jacobi CG of feec_csr_cg on the scratch array above instead of
// allocated buffers: the same kernels in the same order, so the same bits.
feec_status feec_csr_cg_noalloc(feec_csr* a, const double* b, double* x, double rtol,
                                int64_t max_iters, feec_report* report) {
  return guard([&] {
    FEEC_REQUIRE(a && b && x, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(a->nrows == a->ncols, FEEC_E_ARG, "CG needs a square matrix");
    FEEC_REQUIRE(rtol >= 0.0 && max_iters >= 0, FEEC_E_ARG, "rtol >= 0 and max_iters >= 0");
    FEEC_REQUIRE(a->nrows <= kNoAllocMaxRows, FEEC_E_ARG,
                 "feec_csr_cg_noalloc takes at most 262144 rows (its fixed scratch); "
                 "feec_csr_cg has no such limit");
    Ctx* ctx = a->ctx;
    FEEC_HIP(hipSetDevice(ctx->device));
    double* base = scratch_for(ctx->device);
    FEEC_REQUIRE(base, FEEC_E_DEVICE, "scratch not set up on this device");
    const int64_t n = a->nrows;
    double* part = base + kScalarDoubles;
    double* v = part + 2 * NBLK;
    double* db = v;
    double* dx = v + kNoAllocMaxRows;
    double* dinv = v + 2 * kNoAllocMaxRows;
    const CgWork w{v + 3 * kNoAllocMaxRows, v + 4 * kNoAllocMaxRows, v + 5 * kNoAllocMaxRows,
                   v + 6 * kNoAllocMaxRows, part, reinterpret_cast<CgS*>(base)};
    copy_to_device(ctx, db, b, n * 8);
    // Jacobi (precond.rs:53-123): the inverse diagonal, 1 where it is zero
    csr_diagonal(*a, dinv);
    k_reciprocal_or_one<<<(unsigned)std::max<int64_t>(1, (n + 255) / 256), 256, 0, ctx->stream>>>(
        dinv, n, dinv);
    FEEC_HIP(hipGetLastError());
    KrylovStats ks = cg_solve_ws(ctx, *a, dinv, db, dx, rtol, max_iters, w);
    copy_to_host(ctx, x, dx, n * 8);
    if (report) {
      *report = feec_report{};
      report->iters = ks.iters;
      report->converged = ks.converged ? 1 : 0;
      report->residual = ks.precond_residual;
      report->precond_residual = ks.precond_residual;
    }
    FEEC_REQUIRE(ks.converged, FEEC_E_NOT_CONVERGED, "CG did not reach rtol within max_iters");
  });
}

}  // extern "C"
