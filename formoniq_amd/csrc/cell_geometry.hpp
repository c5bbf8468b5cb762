// Reference-cell tables and the per-cell Regge metric, shared by the element
// kernels (elmat.hip) and the quadrature kernels (analytic.hip).
//
// Reference behaviour (luiswirth/formoniq, crates/):
//   SimplexLengthsSq::metric       regge/src/lengths/simplex.rs:308-326
//   Metric::{det, dual}            metric/src/lib.rs:179-190, 250-257
//   determinant (Leibniz)          multialgebra/src/lib.rs:447-456,
//                                  multiindex/src/permutation.rs:160-169
//   unit_difbarys (P = Lambda^K(Dlambda))  simplicial/src/atlas.rs
//   Combination::deletions         multiindex (the deletion pairs of a DOF)
#pragma once

#include "feec_internal.hpp"

namespace feec {
namespace {

constexpr double factorial(int n) { return n <= 1 ? 1.0 : n * factorial(n - 1); }

// The permutations of {0..N-1} in Permutation::all's order: a word runs
// through S_N lexicographically and the permutation is the word reversed.
template <int N>
struct Perms {
  static constexpr int count = N <= 1 ? 1 : (N == 2 ? 2 : (N == 3 ? 6 : 24));
  int p[count][N > 0 ? N : 1];
  bool odd[count];
  constexpr Perms() : p(), odd() {
    int word[N > 0 ? N : 1] = {};
    for (int i = 0; i < N; ++i) word[i] = i;
    for (int c = 0; c < count; ++c) {
      int inv = 0;
      for (int i = 0; i < N; ++i) p[c][i] = word[N - 1 - i];
      for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j)
          if (p[c][i] > p[c][j]) ++inv;
      odd[c] = (inv & 1) != 0;
      // lexicographic successor of the word
      int piv = -1;
      for (int i = N - 2; i >= 0; --i)
        if (word[i] < word[i + 1]) {
          piv = i;
          break;
        }
      if (piv < 0) break;
      int succ = N - 1;
      while (word[succ] <= word[piv]) --succ;
      int t = word[piv];
      word[piv] = word[succ];
      word[succ] = t;
      for (int i = piv + 1, j = N - 1; i < j; ++i, --j) {
        t = word[i];
        word[i] = word[j];
        word[j] = t;
      }
    }
  }
};

// Leibniz determinant, products folded left to right, terms summed in order.
template <int N>
__device__ __forceinline__ double leibniz(const double (&a)[N > 0 ? N : 1][N > 0 ? N : 1]) {
#pragma clang fp contract(off)
  if constexpr (N == 0) {
    return 1.0;
  } else {
    constexpr Perms<N> P{};
    double total = 0.0;
#pragma unroll
    for (int c = 0; c < Perms<N>::count; ++c) {
      double prod = a[0][P.p[c][0]];
#pragma unroll
      for (int i = 1; i < N; ++i) prod = prod * a[i][P.p[c][i]];
      if (c == 0)
        total = P.odd[c] ? -prod : prod;
      else
        total = P.odd[c] ? total - prod : total + prod;
    }
    return total;
  }
}

template <int D>
struct CellGeom {
  double g[D][D];
  double ginv[D][D];
  double det;
};

__host__ __device__ constexpr int edge_index(int i, int j) { return i + j * (j - 1) / 2; }

template <int D>
__device__ __forceinline__ void cell_geometry(const double* s, CellGeom<D>& cg) {
#pragma clang fp contract(off)
#pragma unroll
  for (int i = 0; i < D; ++i) cg.g[i][i] = s[edge_index(0, i + 1)];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = i + 1; j < D; ++j) {
      const double v = 0.5 * ((s[edge_index(0, i + 1)] + s[edge_index(0, j + 1)]) -
                              s[edge_index(i + 1, j + 1)]);
      cg.g[i][j] = v;
      cg.g[j][i] = v;
    }
  cg.det = leibniz<D>(cg.g);
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      double m[D > 1 ? D - 1 : 1][D > 1 ? D - 1 : 1];
#pragma unroll
      for (int a = 0; a < D - 1; ++a)
#pragma unroll
        for (int b = 0; b < D - 1; ++b) m[a][b] = cg.g[a < i ? a : a + 1][b < j ? b : b + 1];
      const double cof = leibniz<D - 1>(m);
      cg.ginv[j][i] = (((i + j) & 1) ? -cof : cof) / cg.det;
    }
}

// Compile-time index tables of grade K on the d-simplex.
constexpr int popc(unsigned m) {
  int c = 0;
  while (m) {
    c += m & 1u;
    m >>= 1;
  }
  return c;
}
constexpr int nth_elem(unsigned m, int p) {
  for (int e = 0; e < 32; ++e)
    if (m & (1u << e)) {
      if (p == 0) return e;
      --p;
    }
  return -1;
}
// position of mask m among the size-popc(m) subsets of {0..n-1} in colex order
constexpr int colex_pos(unsigned m) {
  int r = 0, i = 0;
  for (int e = 0; e < 32; ++e)
    if (m & (1u << e)) {
      r += binom(e, i + 1);
      ++i;
    }
  return r;
}
// k-subsets of {0..n-1}, colex order = increasing mask
constexpr unsigned nth_subset(int n, int k, int idx) {
  for (unsigned m = 0; m < (1u << n); ++m)
    if (popc(m) == k && idx-- == 0) return m;
  return 0;
}
// integer determinant of a {0, +-1} K x K matrix (exact, any order)
template <int K>
constexpr int int_det(const int (&a)[4][4]) {
  if constexpr (K == 0) {
    return 1;
  } else {
    constexpr Perms<K> P{};
    int t = 0;
    for (int c = 0; c < Perms<K>::count; ++c) {
      int prod = 1;
      for (int i = 0; i < K; ++i) prod *= a[i][P.p[c][i]];
      t += P.odd[c] ? -prod : prod;
    }
    return t;
  }
}

template <int D, int K>
struct Grade {
  static constexpr int NA = binom(D, K);          // Lambda^K g^-1 is NA x NA
  static constexpr int NP = binom(D + 1, K);      // H is NP x NP (vertex K-subsets)
  static constexpr int ND = binom(D + 1, K + 1);  // Whitney DOFs
  int ax[NA][K > 0 ? K : 1];    // axis subsets
  int P[NP][NA];                // Lambda^K(Dlambda)
  int dv[ND][K + 1];            // vertex deleted at position p of DOF i
  int dr[ND][K + 1];            // rank of DOF i without it (an H index)
  int bd[NP][ND];               // local boundary d_K: faces x DOFs
  constexpr Grade() : ax(), P(), dv(), dr(), bd() {
    for (int I = 0; I < NA; ++I) {
      unsigned m = nth_subset(D, K, I);
      for (int a = 0; a < K; ++a) ax[I][a] = nth_elem(m, a);
    }
    for (int A = 0; A < NP; ++A) {
      unsigned ma = nth_subset(D + 1, K, A);
      for (int I = 0; I < NA; ++I) {
        unsigned mi = nth_subset(D, K, I);
        int minor[4][4] = {};
        for (int a = 0; a < K; ++a)
          for (int b = 0; b < K; ++b) {
            const int row = nth_elem(ma, a), col = nth_elem(mi, b);
            minor[a][b] = row == 0 ? -1 : (row - 1 == col ? 1 : 0);
          }
        P[A][I] = int_det<K>(minor);
      }
    }
    for (int i = 0; i < ND; ++i) {
      unsigned m = nth_subset(D + 1, K + 1, i);
      for (int p = 0; p <= K; ++p) {
        const int v = nth_elem(m, p);
        dv[i][p] = v;
        dr[i][p] = colex_pos(m & ~(1u << v));
        if (K >= 1) bd[dr[i][p]][i] = (p & 1) ? -1 : 1;
      }
    }
  }
};

}  // namespace
}  // namespace feec
