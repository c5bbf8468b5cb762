// Per-cell Regge metric and Whitney element matrices, FP64, one lane per cell.
//
// Reference behaviour (luiswirth/formoniq, crates/):
//   SimplexLengthsSq::metric       regge/src/lengths/simplex.rs:308-326
//   Metric::{det, dual}            metric/src/lib.rs:179-190, 250-257
//   cell_volume                    regge/src/lib.rs:26-28, simplicial/src/atlas.rs:102-105
//   multiform_metric               metric/src/tensor.rs:127-131
//   determinant (Leibniz)          multialgebra/src/lib.rs:447-456,
//                                  multiindex/src/permutation.rs:160-169
//   HodgeMass::element             formoniq/src/operators.rs:84-94
//   WhitneyExpansion::pullback     derham/src/interpolate/form.rs:195-232
//   unit_bary_gramian              simplicial/src/atlas.rs:198-204
//   WhitneyPairing::element        formoniq/src/operators.rs:201-211
//
// Arithmetic. Every element entry is evaluated with the reference's own
// operations in the reference's order, with FP contraction off, so the
// matrices are bit-identical to the CPU restatement of that order and an entry
// that the reference's arithmetic leaves at a roundoff residue (an
// exactly-representable cell) is nonzero here too: the assembled sparsity
// pattern (`push if v != 0.0`, galerkin.rs:173) is the reference's. Per cell:
//   g       polarization, 0.5 * ((s0i + s0j) - sij);
//   det g   the Leibniz sum over the permutations in colex order, from 0.0;
//   g^-1    adjugate: (+-Leibniz cofactor) / det;
//   L       Lambda^K g^-1, entry (I, J) the Leibniz determinant of the K x K
//           minor of g^-1 on axis subsets I, J (colex);
//   H       (P L) P^T with P = Lambda^K(Dlambda) in {0, +-1}: natural-order
//           sums over the nonzero terms (a product by +-1 is exact);
//   M       vol * sum_{p,q} ((a_p b_q) H[..]) Q[..] over the deletion pairs of
//           the two DOFs, p-major, a_p = (-1)^p K!, Q = (1 + delta)/((d+1)(d+2));
//   dif     M d^T, d M, d (M d^T): natural-order sums of +-1 multiples.
// Products by 0 and +-1 change no bits and the sign of a zero never changes a
// nonzero result, so the sums skip the zero terms and start at their first.
//
// Kernel shape: one wavefront per group of 64 cells, one lane per cell; all
// index tables are compile-time constants per (d, K), so the whole evaluation
// is straight-line register code. Element matrices are written
// structure-of-arrays, out[entry * ncells + cell], so each store instruction of
// a wavefront covers 64 consecutive doubles.
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "cell_geometry.hpp"
#include "feec_internal.hpp"

namespace feec {

namespace {

constexpr int TPB = 256;
constexpr int WAVE = 64;

// Block kinds: 0 mass, 1 dif_trial, 2 dif_test, 3 dif_both, 4 diag(mass), 5 diag(dif_both)
constexpr int KIND_DIAG_MASS = 4, KIND_DIAG_BOTH = 5;
constexpr unsigned B_MASS = 1u << 0, B_TRIAL = 1u << 1, B_TEST = 1u << 2, B_BOTH = 1u << 3,
                   B_DMASS = 1u << 4, B_DBOTH = 1u << 5;

__host__ __device__ constexpr int block_rows(int d, int kind, int K) {
  return (kind == 2 || kind == 3 || kind == KIND_DIAG_BOTH) ? binom(d + 1, K) : binom(d + 1, K + 1);
}
__host__ __device__ constexpr int block_cols(int d, int kind, int K) {
  return (kind == KIND_DIAG_MASS || kind == KIND_DIAG_BOTH) ? 1
         : (kind == 1 || kind == 3)                           ? binom(d + 1, K)
                                                              : binom(d + 1, K + 1);
}

// Output arrays of one grade, by block kind (SoA [entry][ncells]).
struct Outs {
  double* p[6];
};

// The index tables of grade K for the runtime row loop, in constant memory:
// the deletion pairs of every DOF and the values of P and of d_K as doubles.
template <int D, int K>
struct RowTables {
  int dr[Grade<D, K>::ND][K + 1];
  int dv[Grade<D, K>::ND][K + 1];
  double P[Grade<D, K>::NP][Grade<D, K>::NA];
  double bd[Grade<D, K>::NP][Grade<D, K>::ND];
  constexpr RowTables() : dr(), dv(), P(), bd() {
    constexpr Grade<D, K> T{};
    for (int i = 0; i < Grade<D, K>::ND; ++i)
      for (int p = 0; p <= K; ++p) {
        dr[i][p] = T.dr[i][p];
        dv[i][p] = T.dv[i][p];
      }
    for (int A = 0; A < Grade<D, K>::NP; ++A)
      for (int I = 0; I < Grade<D, K>::NA; ++I) P[A][I] = T.P[A][I];
    for (int f = 0; f < Grade<D, K>::NP; ++f)
      for (int i = 0; i < Grade<D, K>::ND; ++i) bd[f][i] = T.bd[f][i];
  }
};
template <int D, int K>
__constant__ RowTables<D, K> kRows = RowTables<D, K>{};

// first l with s[l] != 0 (N if none)
template <int N>
constexpr int first_nz(const int (&s)[N]) {
  for (int l = 0; l < N; ++l)
    if (s[l]) return l;
  return N;
}

// sum_l x[l] * s[l] over s in {0, +-1}, in order of l, from the first nonzero term
template <int N, class F>
__device__ __forceinline__ double signed_sum(const int (&s)[N], F x) {
#pragma clang fp contract(off)
  double acc = 0.0;
  bool first = true;
#pragma unroll
  for (int l = 0; l < N; ++l) {
    if (s[l] == 0) continue;
    const double v = x(l);
    if (first)
      acc = s[l] > 0 ? v : -v;
    else
      acc = s[l] > 0 ? acc + v : acc - v;
    first = false;
  }
  return acc;
}

template <int K>
__host__ __device__ constexpr double pair_weight(int p, int q) {
  // a_p b_q = (+-K!)(+-K!), exact
  return ((p & 1) ? -factorial(K) : factorial(K)) * ((q & 1) ? -factorial(K) : factorial(K));
}

__device__ __forceinline__ void put(double* __restrict__ out, int64_t entry, int64_t nc, int64_t cell,
                                    bool active, double v) {
  // element values are written once and read later by another kernel: a
  // non-temporal store streams them to HBM without allocating in the caches
  if (active) __builtin_nontemporal_store(v, out + entry * nc + cell);
}

// The blocks KINDS of grade K for one cell. M is produced one row at a time
// (row i of the Whitney mass, all columns); rows are emitted as they come and
// the d-products accumulate over them in the reference's summation order:
//   dif_test  (d M)[f][j]      = sum_l d[f][l] M[l][j]   (l = row order)
//   dif_trial (M d^T)[i][g]    = sum_t M[i][t] d[g][t]   (within the row)
//   dif_both  (d (M d^T))[f][g] = sum_l d[f][l] (M d^T)[l][g]
// Small grades (H at most 6 x 6) are fully unrolled with H in registers; the
// 10 x 10 H of the 4-simplex's middle grades is streamed: a runtime loop over
// rows computes, for each deletion p of the row's DOF in order, the one H row
// it needs and adds its terms, so that only that row is live.
template <int D, int K, unsigned KINDS>
__device__ __forceinline__ void grade_blocks(const double (&gi)[D][D], double vol, const Outs& o,
                                             int64_t nc, int64_t cell, bool active) {
#pragma clang fp contract(off)
  using G = Grade<D, K>;
  constexpr G T{};
  constexpr int NA = G::NA, NP = G::NP, ND = G::ND;
  constexpr double qscale = 1.0 / (double)((D + 1) * (D + 2));
  constexpr double q2 = 2.0 * qscale;
  constexpr bool TEST = (KINDS & B_TEST) != 0;
  constexpr bool BOTH = (KINDS & (B_BOTH | B_DBOTH)) != 0;
  constexpr bool MD = (KINDS & (B_TRIAL | B_BOTH | B_DBOTH)) != 0;
  static_assert(K >= 1 || (KINDS & ~(B_MASS | B_DMASS)) == 0, "grade 0 has no d blocks");
  // Lambda^K g^-1
  double L[NA][NA];
#pragma unroll
  for (int I = 0; I < NA; ++I)
#pragma unroll
    for (int J = 0; J < NA; ++J) {
      double m[K > 0 ? K : 1][K > 0 ? K : 1];
#pragma unroll
      for (int a = 0; a < K; ++a)
#pragma unroll
        for (int b = 0; b < K; ++b) m[a][b] = gi[T.ax[I][a]][T.ax[J][b]];
      L[I][J] = leibniz<K>(m);
    }
  double tacc[TEST ? NP : 1][TEST ? ND : 1];
  double bacc[BOTH ? NP : 1][BOTH ? NP : 1];
  constexpr bool STREAM = NP > 6;
  if constexpr (!STREAM) {
    // H = (P L) P^T
    double H[NP][NP];
#pragma unroll
    for (int A = 0; A < NP; ++A) {
      double pl[NA];
#pragma unroll
      for (int J = 0; J < NA; ++J) pl[J] = signed_sum<NA>(T.P[A], [&](int I) { return L[I][J]; });
#pragma unroll
      for (int B = 0; B < NP; ++B) H[A][B] = signed_sum<NA>(T.P[B], [&](int J) { return pl[J]; });
    }
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      double mrow[ND];
#pragma unroll
      for (int j = 0; j < ND; ++j) {
        double sum = 0.0;
#pragma unroll
        for (int p = 0; p <= K; ++p)
#pragma unroll
          for (int q = 0; q <= K; ++q) {
            const double t = (pair_weight<K>(p, q) * H[T.dr[i][p]][T.dr[j][q]]) *
                             (T.dv[i][p] == T.dv[j][q] ? q2 : qscale);
            sum = (p == 0 && q == 0) ? t : sum + t;
          }
        mrow[j] = vol * sum;
      }
      if constexpr ((KINDS & B_MASS) != 0) {
#pragma unroll
        for (int j = 0; j < ND; ++j) put(o.p[0], i * ND + j, nc, cell, active, mrow[j]);
      }
      if constexpr ((KINDS & B_DMASS) != 0) put(o.p[KIND_DIAG_MASS], i, nc, cell, active, mrow[i]);
      if constexpr (TEST) {
#pragma unroll
        for (int f = 0; f < NP; ++f) {
          if (T.bd[f][i] == 0) continue;
#pragma unroll
          for (int j = 0; j < ND; ++j) {
            const double v = T.bd[f][i] > 0 ? mrow[j] : -mrow[j];
            tacc[f][j] = (first_nz(T.bd[f]) == i) ? v : tacc[f][j] + v;
          }
        }
      }
      if constexpr (MD) {
        double mdrow[NP];
#pragma unroll
        for (int g = 0; g < NP; ++g) mdrow[g] = signed_sum<ND>(T.bd[g], [&](int t) { return mrow[t]; });
        if constexpr ((KINDS & B_TRIAL) != 0) {
#pragma unroll
          for (int g = 0; g < NP; ++g) put(o.p[1], i * NP + g, nc, cell, active, mdrow[g]);
        }
        if constexpr (BOTH) {
#pragma unroll
          for (int f = 0; f < NP; ++f) {
            if (T.bd[f][i] == 0) continue;
#pragma unroll
            for (int g = 0; g < NP; ++g) {
              if ((KINDS & B_BOTH) == 0 && g != f) continue;
              const double v = T.bd[f][i] > 0 ? mdrow[g] : -mdrow[g];
              bacc[f][g] = (first_nz(T.bd[f]) == i) ? v : bacc[f][g] + v;
            }
          }
        }
      }
    }
  } else {
    const RowTables<D, K>& R = kRows<D, K>;
    if constexpr (TEST) {
#pragma unroll
      for (int f = 0; f < NP; ++f)
#pragma unroll
        for (int j = 0; j < ND; ++j) tacc[f][j] = 0.0;
    }
    if constexpr (BOTH) {
#pragma unroll
      for (int f = 0; f < NP; ++f)
#pragma unroll
        for (int g = 0; g < NP; ++g) bacc[f][g] = 0.0;
    }
#pragma unroll 1
    for (int i = 0; i < ND; ++i) {
      double mrow[ND];
#pragma unroll
      for (int p = 0; p <= K; ++p) {
        const int A = R.dr[i][p], va = R.dv[i][p];
        // row A of P L: all NA terms from 0.0 (a product by 0 or +-1 is exact)
        double pl[NA];
#pragma unroll
        for (int J = 0; J < NA; ++J) {
          double acc = 0.0;
#pragma unroll
          for (int I = 0; I < NA; ++I) acc = acc + R.P[A][I] * L[I][J];
          pl[J] = acc;
        }
        double h[NP];
#pragma unroll
        for (int B = 0; B < NP; ++B) h[B] = signed_sum<NA>(T.P[B], [&](int J) { return pl[J]; });
#pragma unroll
        for (int j = 0; j < ND; ++j)
#pragma unroll
          for (int q = 0; q <= K; ++q) {
            const double t = (pair_weight<K>(p, q) * h[T.dr[j][q]]) * (va == T.dv[j][q] ? q2 : qscale);
            mrow[j] = (p == 0 && q == 0) ? t : mrow[j] + t;
          }
      }
#pragma unroll
      for (int j = 0; j < ND; ++j) mrow[j] = vol * mrow[j];
      if constexpr ((KINDS & B_MASS) != 0) {
#pragma unroll
        for (int j = 0; j < ND; ++j) put(o.p[0], i * ND + j, nc, cell, active, mrow[j]);
      }
      if constexpr ((KINDS & B_DMASS) != 0) {
#pragma unroll
        for (int j = 0; j < ND; ++j)
          if (j == i) put(o.p[KIND_DIAG_MASS], i, nc, cell, active, mrow[j]);
      }
      if constexpr (TEST) {
#pragma unroll
        for (int f = 0; f < NP; ++f) {
          const double s = R.bd[f][i];
#pragma unroll
          for (int j = 0; j < ND; ++j) tacc[f][j] = tacc[f][j] + s * mrow[j];
        }
      }
      if constexpr (MD) {
        double mdrow[NP];
#pragma unroll
        for (int g = 0; g < NP; ++g) mdrow[g] = signed_sum<ND>(T.bd[g], [&](int t) { return mrow[t]; });
        if constexpr ((KINDS & B_TRIAL) != 0) {
#pragma unroll
          for (int g = 0; g < NP; ++g) put(o.p[1], i * NP + g, nc, cell, active, mdrow[g]);
        }
        if constexpr (BOTH) {
#pragma unroll
          for (int f = 0; f < NP; ++f) {
            const double s = R.bd[f][i];
#pragma unroll
            for (int g = 0; g < NP; ++g)
              if ((KINDS & B_BOTH) != 0 || g == f) bacc[f][g] = bacc[f][g] + s * mdrow[g];
          }
        }
      }
    }
  }
  if constexpr (TEST) {
#pragma unroll
    for (int f = 0; f < NP; ++f)
#pragma unroll
      for (int j = 0; j < ND; ++j) put(o.p[2], f * ND + j, nc, cell, active, tacc[f][j]);
  }
  if constexpr ((KINDS & B_BOTH) != 0) {
#pragma unroll
    for (int f = 0; f < NP; ++f)
#pragma unroll
      for (int g = 0; g < NP; ++g) put(o.p[3], f * NP + g, nc, cell, active, bacc[f][g]);
  }
  if constexpr ((KINDS & B_DBOTH) != 0) {
#pragma unroll
    for (int f = 0; f < NP; ++f) put(o.p[KIND_DIAG_BOTH], f, nc, cell, active, bacc[f][f]);
  }
}

// Per-cell set-up shared by every job: lengths, g, det g, g^-1, vol.
template <int D>
struct CellIn {
  CellGeom<D> cg;
  double vol;
  bool active;
  int64_t cell;
};

template <int D>
__device__ __forceinline__ void cell_in(const int32_t* __restrict__ edges, const double* __restrict__ len,
                                        int64_t nc, int64_t cell0, int* degenerate, CellIn<D>& c) {
#pragma clang fp contract(off)
  constexpr int NE = binom(D + 1, 2);
  c.cell = cell0 + threadIdx.x;
  c.active = c.cell < nc;
  double s[NE];
  // inactive lanes evaluate the unit right-angled simplex (nonsingular)
#pragma unroll
  for (int e = 0; e < NE; ++e) s[e] = c.active ? len[edges[c.cell * NE + e]] : ((e < D) ? 1.0 : 2.0);
  cell_geometry<D>(s, c.cg);
  if (c.active && c.cg.det == 0.0) atomicOr(degenerate, 1);
  constexpr double unit_vol = 1.0 / factorial(D);
  c.vol = unit_vol * sqrt(fabs(c.cg.det));
}

// Jobs. KKT of grade k: M_{k-1}; dif_test(k), M_k, diag(M_k), dif_both(k) and
// its diagonal; dif_both(k+1). Dirac: M_k (all k) and dif_trial(k) (k >= 1).
// Single: one block (K, kind) chosen at run time.
struct KktOuts {
  Outs g[3];  // grades k-1, k, k+1
};
struct DiracOuts {
  Outs g[5];
};

template <int D, int k>
__device__ __forceinline__ void kkt_cell(const CellIn<D>& c, const KktOuts& o, int64_t nc) {
  if constexpr (k >= 1) grade_blocks<D, k - 1, B_MASS>(c.cg.ginv, c.vol, o.g[0], nc, c.cell, c.active);
  if constexpr (k >= 1)
    grade_blocks<D, k, B_MASS | B_DMASS | B_TEST | B_BOTH | B_DBOTH>(c.cg.ginv, c.vol, o.g[1], nc,
                                                                     c.cell, c.active);
  else
    grade_blocks<D, k, B_MASS | B_DMASS>(c.cg.ginv, c.vol, o.g[1], nc, c.cell, c.active);
  if constexpr (k + 1 <= D) grade_blocks<D, k + 1, B_BOTH>(c.cg.ginv, c.vol, o.g[2], nc, c.cell, c.active);
}

template <int D, int K = 0>
__device__ __forceinline__ void dirac_cell(const CellIn<D>& c, const DiracOuts& o, int64_t nc) {
  if constexpr (K <= D) {
    if constexpr (K == 0)
      grade_blocks<D, 0, B_MASS>(c.cg.ginv, c.vol, o.g[0], nc, c.cell, c.active);
    else
      grade_blocks<D, K, B_MASS | B_TRIAL>(c.cg.ginv, c.vol, o.g[K], nc, c.cell, c.active);
    dirac_cell<D, K + 1>(c, o, nc);
  }
}

template <int D, int K = 0>
__device__ __forceinline__ void single_cell(const CellIn<D>& c, int grade, int kind, const Outs& o,
                                            int64_t nc) {
  if constexpr (K <= D) {
    if (grade == K) {
      switch (kind) {
        case 0: grade_blocks<D, K, B_MASS>(c.cg.ginv, c.vol, o, nc, c.cell, c.active); return;
        case KIND_DIAG_MASS: grade_blocks<D, K, B_DMASS>(c.cg.ginv, c.vol, o, nc, c.cell, c.active); return;
        default: break;
      }
      if constexpr (K >= 1) {
        switch (kind) {
          case 1: grade_blocks<D, K, B_TRIAL>(c.cg.ginv, c.vol, o, nc, c.cell, c.active); return;
          case 2: grade_blocks<D, K, B_TEST>(c.cg.ginv, c.vol, o, nc, c.cell, c.active); return;
          case 3: grade_blocks<D, K, B_BOTH>(c.cg.ginv, c.vol, o, nc, c.cell, c.active); return;
          case KIND_DIAG_BOTH: grade_blocks<D, K, B_DBOTH>(c.cg.ginv, c.vol, o, nc, c.cell, c.active); return;
          default: break;
        }
      }
      return;
    }
    single_cell<D, K + 1>(c, grade, kind, o, nc);
  }
}

template <int D>
__global__ void __launch_bounds__(WAVE) k_elmat_single(const int32_t* __restrict__ edges,
                                                       const double* __restrict__ len, int64_t nc,
                                                       int grade, int kind, Outs o, int* degenerate) {
  CellIn<D> c;
  cell_in<D>(edges, len, nc, blockIdx.x * (int64_t)WAVE, degenerate, c);
  single_cell<D>(c, grade, kind, o, nc);
}

template <int D, int k>
__global__ void __launch_bounds__(WAVE) k_elmat_kkt(const int32_t* __restrict__ edges,
                                                    const double* __restrict__ len, int64_t nc,
                                                    KktOuts o, int* degenerate) {
  CellIn<D> c;
  cell_in<D>(edges, len, nc, blockIdx.x * (int64_t)WAVE, degenerate, c);
  kkt_cell<D, k>(c, o, nc);
}

template <int D>
__global__ void __launch_bounds__(WAVE) k_elmat_dirac(const int32_t* __restrict__ edges,
                                                      const double* __restrict__ len, int64_t nc,
                                                      DiracOuts o, int* degenerate) {
  CellIn<D> c;
  cell_in<D>(edges, len, nc, blockIdx.x * (int64_t)WAVE, degenerate, c);
  dirac_cell<D>(c, o, nc);
}

// One grade's blocks of a KKT or Dirac launch on its own: the split variant of
// the fused kernels above. Each entry is computed by the same operations as in
// the fused kernel (bit-identical); a lane holds one grade's state only, so
// more wavefronts fit per SIMD, at the cost of recomputing the cell metric
// once per grade.
template <int D, int K, unsigned KINDS>
__global__ void __launch_bounds__(WAVE) k_elmat_grade(const int32_t* __restrict__ edges,
                                                      const double* __restrict__ len, int64_t nc,
                                                      Outs o, int* degenerate) {
  CellIn<D> c;
  cell_in<D>(edges, len, nc, blockIdx.x * (int64_t)WAVE, degenerate, c);
  grade_blocks<D, K, KINDS>(c.cg.ginv, c.vol, o, nc, c.cell, c.active);
}

// Many meshes of mixed dimension and grade in one persistent launch: every
// wavefront takes (mesh, 64-cell range) items from a global counter until the
// queue is empty, so the grid drains whatever the mix of sizes.
struct BatchMesh {
  int D, k;
  int64_t nc;
  const int32_t* edges;
  const double* len;
  int64_t item0;  // first work item of this mesh
  KktOuts outs;
};

template <int D, int k = 0>
__device__ __forceinline__ void batched_group(const BatchMesh& m, int64_t cell0, int* degenerate) {
  if constexpr (k <= D) {
    if (m.k == k) {
      CellIn<D> c;
      cell_in<D>(m.edges, m.len, m.nc, cell0, degenerate, c);
      kkt_cell<D, k>(c, m.outs, m.nc);
      return;
    }
    batched_group<D, k + 1>(m, cell0, degenerate);
  }
}

__global__ void __launch_bounds__(WAVE) k_elmat_batched(const BatchMesh* __restrict__ meshes,
                                                        int nmeshes, int64_t nitems,
                                                        unsigned long long* counter,
                                                        int* degenerate) {
  __shared__ long long item;
  for (;;) {
    if (threadIdx.x == 0) item = (long long)atomicAdd(counter, 1ull);
    __syncthreads();
    const long long it = item;
    __syncthreads();
    if (it >= nitems) break;
    int lo = 0, hi = nmeshes - 1;  // last mesh with item0 <= it
    while (lo < hi) {
      int mid = (lo + hi + 1) / 2;
      if (meshes[mid].item0 <= it) lo = mid; else hi = mid - 1;
    }
    const BatchMesh& m = meshes[lo];
    const int64_t cell0 = (it - m.item0) * WAVE;
    switch (m.D) {
      case 2: batched_group<2>(m, cell0, degenerate); break;
      case 3: batched_group<3>(m, cell0, degenerate); break;
      case 4: batched_group<4>(m, cell0, degenerate); break;
      default: break;
    }
  }
}
template <int D>
__global__ void k_metrics(const int32_t* __restrict__ edges, const double* __restrict__ len,
                          int64_t nc, double* g, double* ginv, double* det) {
  int64_t cell = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (cell >= nc) return;
  constexpr int NE = binom(D + 1, 2);
  double s[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) s[e] = len[edges[cell * NE + e]];
  CellGeom<D> cg;
  cell_geometry<D>(s, cg);
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      if (g) g[cell * D * D + i * D + j] = cg.g[i][j];
      if (ginv) ginv[cell * D * D + i * D + j] = cg.ginv[i][j];
    }
  if (det) det[cell] = cg.det;
}

// Riemannian iff the Cholesky factorization of g succeeds (metric/src/lib.rs:203-205).
template <int D>
__global__ void k_riemannian(const int32_t* __restrict__ edges, const double* __restrict__ len,
                             int64_t nc, uint8_t* riem) {
  int64_t cell = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (cell >= nc) return;
  constexpr int NE = binom(D + 1, 2);
  double s[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) s[e] = len[edges[cell * NE + e]];
  CellGeom<D> cg;
  cell_geometry<D>(s, cg);
  double L[D][D] = {};
  bool ok = true;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    double dsum = cg.g[j][j];
#pragma unroll
    for (int k = 0; k < j; ++k) dsum -= L[j][k] * L[j][k];
    if (!(dsum > 0.0)) ok = false;
    double ljj = sqrt(fmax(dsum, 0.0));
    L[j][j] = ljj;
#pragma unroll
    for (int i = j + 1; i < D; ++i) {
      double v = cg.g[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
      L[i][j] = ljj > 0.0 ? v / ljj : 0.0;
    }
  }
  riem[cell] = ok ? 1 : 0;
}

__global__ void k_census(const double* len, int64_t ne, unsigned long long* counts) {
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double s = len[e];
  // one atomic per wavefront and category (lane 0 is active whenever any lane is)
  const unsigned long long neg = __ballot(s < 0.0), zero = __ballot(s == 0.0),
                           pos = __ballot(!(s < 0.0) && !(s == 0.0));
  if ((threadIdx.x & 63) == 0) {
    if (neg) atomicAdd(&counts[0], (unsigned long long)__popcll(neg));
    if (zero) atomicAdd(&counts[1], (unsigned long long)__popcll(zero));
    if (pos) atomicAdd(&counts[2], (unsigned long long)__popcll(pos));
  }
}

inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + TPB - 1) / TPB); }

void check_degenerate(Ctx* ctx, int* deg) {
  int h = 0;
  FEEC_HIP(hipMemcpyAsync(&h, deg, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  FEEC_REQUIRE(h == 0, FEEC_E_DEGENERATE_CELL, "a cell metric is degenerate (det g == 0)");
}

}  // namespace

int element_rows(int d, int kind, int grade) {
  bool test_dif = (kind == 2 || kind == 3);
  return test_dif ? binom(d + 1, grade) : binom(d + 1, grade + 1);
}
int element_cols(int d, int kind, int grade) {
  bool trial_dif = (kind == 1 || kind == 3);
  return trial_dif ? binom(d + 1, grade) : binom(d + 1, grade + 1);
}

namespace {

unsigned wave_grid(int64_t nc) { return (unsigned)std::max<int64_t>(1, (nc + WAVE - 1) / WAVE); }

void reset_degenerate(Ctx* ctx) { FEEC_HIP(hipMemsetAsync(ctx->dflag, 0, sizeof(int), ctx->stream)); }

KktOuts kkt_outs(int k, double* const* outs) {
  KktOuts o{};
  if (k >= 1) o.g[0].p[0] = outs[0];
  o.g[1].p[2] = outs[1];
  o.g[1].p[KIND_DIAG_BOTH] = outs[3];
  o.g[1].p[KIND_DIAG_MASS] = outs[4];
  o.g[1].p[0] = outs[5];
  o.g[1].p[3] = outs[6];
  o.g[2].p[3] = outs[2];
  return o;
}

// Element-kernel variants, tuned per (d, k, size) by tune_run: 0 fused (one
// launch, every grade's blocks per lane), 1 split (one launch per grade).
constexpr int kElementVariants = 2;

template <int D, int k>
void launch_kkt_variant(Ctx* ctx, const int32_t* edges, const double* len, int64_t nc,
                        const KktOuts& o, int variant) {
  const unsigned grid = wave_grid(nc);
  if (variant == 0) {
    k_elmat_kkt<D, k><<<grid, WAVE, 0, ctx->stream>>>(edges, len, nc, o, ctx->dflag);
    return;
  }
  if constexpr (k >= 1) {
    k_elmat_grade<D, k - 1, B_MASS><<<grid, WAVE, 0, ctx->stream>>>(edges, len, nc, o.g[0], ctx->dflag);
    k_elmat_grade<D, k, B_MASS | B_DMASS | B_TEST | B_BOTH | B_DBOTH>
        <<<grid, WAVE, 0, ctx->stream>>>(edges, len, nc, o.g[1], ctx->dflag);
  } else {
    k_elmat_grade<D, k, B_MASS | B_DMASS><<<grid, WAVE, 0, ctx->stream>>>(edges, len, nc, o.g[1],
                                                                          ctx->dflag);
  }
  if constexpr (k + 1 <= D)
    k_elmat_grade<D, k + 1, B_BOTH><<<grid, WAVE, 0, ctx->stream>>>(edges, len, nc, o.g[2], ctx->dflag);
}

template <int D, int k = 0>
void launch_kkt(Ctx* ctx, const int32_t* edges, const double* len, int64_t nc, int grade,
                const KktOuts& o) {
  if constexpr (k <= D) {
    if (grade == k) {
      tune_run(ctx, "elmat_kkt", D, k, nc, kElementVariants, [&](int v) {
        launch_kkt_variant<D, k>(ctx, edges, len, nc, o, v);
        FEEC_HIP(hipGetLastError());
      });
      return;
    }
    launch_kkt<D, k + 1>(ctx, edges, len, nc, grade, o);
  } else {
    fail(FEEC_E_ARG, "grade out of range for a KKT system");
  }
}

template <int D, int K = 0>
void launch_dirac_split(Ctx* ctx, const int32_t* edges, const double* len, int64_t nc,
                        const DiracOuts& o) {
  if constexpr (K <= D) {
    const unsigned grid = wave_grid(nc);
    if constexpr (K == 0)
      k_elmat_grade<D, 0, B_MASS><<<grid, WAVE, 0, ctx->stream>>>(edges, len, nc, o.g[0], ctx->dflag);
    else
      k_elmat_grade<D, K, B_MASS | B_TRIAL><<<grid, WAVE, 0, ctx->stream>>>(edges, len, nc, o.g[K],
                                                                             ctx->dflag);
    launch_dirac_split<D, K + 1>(ctx, edges, len, nc, o);
  }
}

template <int D>
void launch_dirac(Ctx* ctx, const int32_t* edges, const double* len, int64_t nc, const DiracOuts& o) {
  tune_run(ctx, "elmat_dirac", D, -1, nc, kElementVariants, [&](int v) {
    if (v == 0)
      k_elmat_dirac<D><<<wave_grid(nc), WAVE, 0, ctx->stream>>>(edges, len, nc, o, ctx->dflag);
    else
      launch_dirac_split<D>(ctx, edges, len, nc, o);
    FEEC_HIP(hipGetLastError());
  });
}

}  // namespace

void element_matrices_soa(Geom& g, int kind, int grade, double* out_dev) {
  Mesh& m = *g.mesh;
  Ctx* ctx = m.ctx;
  const int64_t nc = m.n[m.d];
  Outs o{};
  o.p[kind] = out_dev;
  reset_degenerate(ctx);
  const int32_t* edges = m.faces[1].p;
  switch (m.d) {
    case 1: k_elmat_single<1><<<wave_grid(nc), WAVE, 0, ctx->stream>>>(edges, g.lengths.p, nc, grade, kind, o, ctx->dflag); break;
    case 2: k_elmat_single<2><<<wave_grid(nc), WAVE, 0, ctx->stream>>>(edges, g.lengths.p, nc, grade, kind, o, ctx->dflag); break;
    case 3: k_elmat_single<3><<<wave_grid(nc), WAVE, 0, ctx->stream>>>(edges, g.lengths.p, nc, grade, kind, o, ctx->dflag); break;
    case 4: k_elmat_single<4><<<wave_grid(nc), WAVE, 0, ctx->stream>>>(edges, g.lengths.p, nc, grade, kind, o, ctx->dflag); break;
  }
  FEEC_HIP(hipGetLastError());
  check_degenerate(ctx, ctx->dflag);
}

// outs: M_{k-1}, dif_test(k), dif_both(k+1), diag(dif_both(k)), diag(M_k), M_k, dif_both(k)
void kkt_element_blocks(Geom& g, int grade, double* const* outs, bool check) {
  Mesh& m = *g.mesh;
  Ctx* ctx = m.ctx;
  const int64_t nc = m.n[m.d];
  const KktOuts o = kkt_outs(grade, outs);
  reset_degenerate(ctx);
  const int32_t* edges = m.faces[1].p;
  switch (m.d) {
    case 1: launch_kkt<1>(ctx, edges, g.lengths.p, nc, grade, o); break;
    case 2: launch_kkt<2>(ctx, edges, g.lengths.p, nc, grade, o); break;
    case 3: launch_kkt<3>(ctx, edges, g.lengths.p, nc, grade, o); break;
    case 4: launch_kkt<4>(ctx, edges, g.lengths.p, nc, grade, o); break;
  }
  FEEC_HIP(hipGetLastError());
  if (check) check_degenerate(ctx, ctx->dflag);
}

// Element blocks of many KKT problems (dimensions 2..4) in one persistent
// launch on ctx's stream. The grid size is picked once per device by timing a
// few candidates on the first batch (the launch is idempotent).
double kkt_element_blocks_batched(Ctx* ctx, const std::vector<BatchElementRequest>& reqs) {
  if (reqs.empty()) return 0.0;
  std::vector<BatchMesh> host(reqs.size());
  int64_t items = 0;
  for (size_t i = 0; i < reqs.size(); ++i) {
    Mesh& m = *reqs[i].g->mesh;
    FEEC_REQUIRE(m.d >= 2 && m.d <= 4, FEEC_E_UNSUPPORTED, "batched assembly covers dimensions 2..4");
    FEEC_REQUIRE(reqs[i].grade >= 0 && reqs[i].grade <= m.d, FEEC_E_ARG, "grade out of range");
    BatchMesh& b = host[i];
    b.D = m.d;
    b.k = reqs[i].grade;
    b.nc = m.n[m.d];
    b.edges = m.faces[1].p;
    b.len = reqs[i].g->lengths.p;
    b.item0 = items;
    b.outs = kkt_outs(reqs[i].grade, reqs[i].outs);
    items += (b.nc + WAVE - 1) / WAVE;
  }
  DBuf<BatchMesh> dm(ctx, host.size());
  FEEC_HIP(hipMemcpyAsync(dm.p, host.data(), host.size() * sizeof(BatchMesh), hipMemcpyHostToDevice,
                          ctx->stream));
  reset_degenerate(ctx);
  auto launch = [&](unsigned waves) {
    FEEC_HIP(hipMemsetAsync(ctx->dcount, 0, sizeof(unsigned long long), ctx->stream));
    k_elmat_batched<<<waves, WAVE, 0, ctx->stream>>>(dm.p, (int)host.size(), items, ctx->dcount,
                                                     ctx->dflag);
    FEEC_HIP(hipGetLastError());
  };
  // the persistent grid, in wavefronts, tuned per device and work size
  static const unsigned kWaves[] = {1024, 2048, 4096, 8192, 16384};
  const int c = tune_run(ctx, "elmat_batched_waves", -1, -1, items, 5,
                         [&](int i) { launch(kWaves[i]); });
  hipEvent_t e0, e1;
  FEEC_HIP(hipEventCreate(&e0));
  FEEC_HIP(hipEventCreate(&e1));
  float best_ms = 0.f;
  FEEC_HIP(hipEventRecord(e0, ctx->stream));
  launch(kWaves[c]);
  FEEC_HIP(hipEventRecord(e1, ctx->stream));
  FEEC_HIP(hipEventSynchronize(e1));
  FEEC_HIP(hipEventElapsedTime(&best_ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  check_degenerate(ctx, ctx->dflag);
  return best_ms * 1e-3;
}

int dirac_element_count(int d) {
  int total = 0;
  for (int k = 0; k <= d; ++k) {
    total += binom(d + 1, k + 1) * binom(d + 1, k + 1);
    if (k >= 1) total += binom(d + 1, k + 1) * binom(d + 1, k);
  }
  return total;
}

// All-grade Hodge-Dirac element matrices: M_k (k = 0..d) and dif_trial(k)
// (k = 1..d), back to back per cell in entry-major SoA order.
void dirac_element_matrices(Geom& g, double* out_dev, bool check) {
  Mesh& m = *g.mesh;
  Ctx* ctx = m.ctx;
  const int d = m.d;
  const int64_t nc = m.n[d];
  DiracOuts o{};
  int64_t off = 0;
  for (int k = 0; k <= d; ++k) {
    o.g[k].p[0] = out_dev + off * nc;
    off += block_rows(d, 0, k) * block_cols(d, 0, k);
    if (k >= 1) {
      o.g[k].p[1] = out_dev + off * nc;
      off += block_rows(d, 1, k) * block_cols(d, 1, k);
    }
  }
  reset_degenerate(ctx);
  const int32_t* edges = m.faces[1].p;
  switch (d) {
    case 1: launch_dirac<1>(ctx, edges, g.lengths.p, nc, o); break;
    case 2: launch_dirac<2>(ctx, edges, g.lengths.p, nc, o); break;
    case 3: launch_dirac<3>(ctx, edges, g.lengths.p, nc, o); break;
    case 4: launch_dirac<4>(ctx, edges, g.lengths.p, nc, o); break;
  }
  FEEC_HIP(hipGetLastError());
  if (check) check_degenerate(ctx, ctx->dflag);
}

void cell_metrics(Geom& g, double* gmat, double* ginv, double* det) {
  Mesh& m = *g.mesh;
  Ctx* ctx = m.ctx;
  int64_t nc = m.n[m.d];
  const int32_t* edges = m.faces[1].p;
  switch (m.d) {
    case 1: k_metrics<1><<<grid_for(nc), TPB, 0, ctx->stream>>>(edges, g.lengths.p, nc, gmat, ginv, det); break;
    case 2: k_metrics<2><<<grid_for(nc), TPB, 0, ctx->stream>>>(edges, g.lengths.p, nc, gmat, ginv, det); break;
    case 3: k_metrics<3><<<grid_for(nc), TPB, 0, ctx->stream>>>(edges, g.lengths.p, nc, gmat, ginv, det); break;
    case 4: k_metrics<4><<<grid_for(nc), TPB, 0, ctx->stream>>>(edges, g.lengths.p, nc, gmat, ginv, det); break;
  }
  FEEC_HIP(hipGetLastError());
}

void causal_census(Geom& g, int64_t counts[3], uint8_t* riem_dev) {
  Mesh& m = *g.mesh;
  Ctx* ctx = m.ctx;
  DBuf<unsigned long long> c(ctx, 3);
  FEEC_HIP(hipMemsetAsync(c.p, 0, 3 * sizeof(unsigned long long), ctx->stream));
  k_census<<<grid_for(m.n[1]), TPB, 0, ctx->stream>>>(g.lengths.p, m.n[1], c.p);
  if (riem_dev) {
    int64_t nc = m.n[m.d];
    const int32_t* edges = m.faces[1].p;
    switch (m.d) {
      case 1: k_riemannian<1><<<grid_for(nc), TPB, 0, ctx->stream>>>(edges, g.lengths.p, nc, riem_dev); break;
      case 2: k_riemannian<2><<<grid_for(nc), TPB, 0, ctx->stream>>>(edges, g.lengths.p, nc, riem_dev); break;
      case 3: k_riemannian<3><<<grid_for(nc), TPB, 0, ctx->stream>>>(edges, g.lengths.p, nc, riem_dev); break;
      case 4: k_riemannian<4><<<grid_for(nc), TPB, 0, ctx->stream>>>(edges, g.lengths.p, nc, riem_dev); break;
    }
  }
  unsigned long long h[3];
  FEEC_HIP(hipMemcpyAsync(h, c.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  for (int i = 0; i < 3; ++i) counts[i] = (int64_t)h[i];
}

}  // namespace feec
