/*
 * feec_oracle.c -- CPU FP64 restatement of formoniq's FEEC assembly-and-solve
 * path (luiswirth/formoniq; paths below are relative to its crates/ folder).
 *
 * TEST INFRASTRUCTURE ONLY. This is the parity checker and the CPU baseline:
 * it is loaded by the tests, the smoke check and the benchmark's CPU leg, never
 * by the product library. It restates the reference algorithm as literally as
 * plain C allows, so that each step can be read side by side with the Rust it
 * follows:
 *   element matrices  formoniq/src/operators.rs:58-212, derham/src/interpolate/form.rs:195-232
 *   metric            regge/src/lengths/simplex.rs:308-326, metric/src/lib.rs:179-257
 *   Lambda^k minors   metric/src/tensor.rs:127-131, multialgebra/src/lib.rs:285-305, 447-514
 *   reference cell    simplicial/src/atlas.rs:102-105, 174-204; topology/simplex.rs:209-238
 *   numbering         simplicial/src/topology/complex.rs:298-336, skeleton.rs:50-83, simplex.rs:126-130
 *   assembly          formoniq/src/galerkin.rs:138-188
 *   Krylov            iterative/src/krylov.rs:48-211, precond.rs:53-166
 * External semantics restated (not vendored in the reference): nalgebra-sparse
 * 0.12 COO->CSR (rows ascending, columns ascending, duplicates summed, summed
 * zeros kept) and CSR x CSR (structural product).
 */
#include "feec_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXD 4
#define MAXV (MAXD + 1)

/* Deliberate mistakes, compiled in only by the test that checks the known
 * answers catch them (0 = the restatement as written). */
#ifndef ORC_MUTATION
#define ORC_MUTATION 0
#endif

/* ------------------------------------------------------------------------ */
/* combinatorics                                                             */
/* ------------------------------------------------------------------------ */

int64_t orc_binom(int64_t n, int64_t k) {
  if (k < 0 || n < 0 || k > n) return 0;
  int64_t r = 1;
  for (int64_t i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

static int popcount(unsigned m) { return __builtin_popcount(m); }

/* Colex (k+1)-subsets of {0..d} as bitmasks: increasing mask value is colex
 * order (the highest element is the most significant), as
 * simplicial/src/topology/simplex.rs:184-193 (unit_subsimps) enumerates. */
static int subset_masks(int n, int size, unsigned *out) {
  int cnt = 0;
  if (size < 0 || size > n) return 0;
  for (unsigned m = 0; m < (1u << n); ++m)
    if (popcount(m) == size) out[cnt++] = m;
  return cnt;
}

/* Colex rank of a subset of {0..}: sum_i C(e_i, i+1)
 * (multiindex/src/combination.rs:107-109). */
static int mask_rank(unsigned m) {
  int r = 0, i = 0;
  for (int e = 0; e < 32; ++e)
    if (m & (1u << e)) {
      r += (int)orc_binom(e, i + 1);
      ++i;
    }
  return r;
}

static int mask_elem(unsigned m, int p) { /* p-th smallest element */
  for (int e = 0; e < 32; ++e)
    if (m & (1u << e)) {
      if (p == 0) return e;
      --p;
    }
  return -1;
}

int orc_local_subsets(int d, int k, int *out) {
  unsigned masks[32];
  int cnt = subset_masks(d + 1, k + 1, masks);
  for (int i = 0; i < cnt; ++i)
    for (int p = 0; p <= k; ++p) out[i * (k + 1) + p] = mask_elem(masks[i], p);
  return cnt;
}

static double factorial(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= (double)i;
  return f;
}

/* Leibniz determinant (multialgebra/src/lib.rs:447-456): the sum, from 0.0,
 * of sign(sigma) * prod_i a[i][sigma(i)] (the product folded from 1.0), over
 * the permutations in the order Permutation::all enumerates them
 * (multiindex/src/permutation.rs:160-169): the lexicographic successor drives a
 * word and the emitted permutation is the word reversed (colex order). For
 * n >= 3 this order decides the rounding of the sum. Row-major n x n, n <= 4. */
static double leibniz_det(const double *a, int n) {
  if (n == 0) return 1.0;
  int word[MAXD], perm[MAXD];
  for (int i = 0; i < n; ++i) word[i] = i;
  double total = 0.0;
  for (;;) {
    for (int i = 0; i < n; ++i) perm[i] = word[n - 1 - i];
    int inv = 0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j)
        if (perm[i] > perm[j]) ++inv;
    double prod = 1.0;
    for (int i = 0; i < n; ++i) prod *= a[i * n + perm[i]];
    total += (inv & 1) ? -prod : prod;
    int piv = -1;
    for (int i = n - 2; i >= 0; --i)
      if (word[i] < word[i + 1]) {
        piv = i;
        break;
      }
    if (piv < 0) break;
    int succ = n - 1;
    while (word[succ] <= word[piv]) --succ;
    int t = word[piv];
    word[piv] = word[succ];
    word[succ] = t;
    for (int i = piv + 1, j = n - 1; i < j; ++i, --j) {
      t = word[i];
      word[i] = word[j];
      word[j] = t;
    }
  }
  return total;
}

/* ------------------------------------------------------------------------ */
/* per-cell geometry                                                         */
/* ------------------------------------------------------------------------ */

/* Local colex edge index of positions i<j: i + j(j-1)/2
 * (simplicial/src/topology/simplex.rs:209-211). */
static int edge_index(int i, int j) { return i + j * (j - 1) / 2; }

/* Polarization g_aa = s(0,a+1), g_ab = (s(0,a+1)+s(0,b+1)-s(a+1,b+1))/2
 * (regge/src/lengths/simplex.rs:308-326). det g and g^-1: the reference calls
 * nalgebra 0.35's determinant and try_inverse (metric/src/lib.rs:179-190,
 * 250-257), closed forms up to 3x3 (4x4: LU / cofactor expansion) whose
 * rounding is not vendored in the reference. Here: the Leibniz sum above and
 * the adjugate, cofactor / det. Parity with nalgebra is unpinned in the last
 * bit; on every cell whose g has small-integer entries times a power of two
 * (the unjittered Kuhn cells) every step is exact under either algorithm. */
void orc_cell_metric(int d, const double *s, double *g, double *ginv, double *det) {
  for (int i = 0; i < d; ++i) g[i * d + i] = s[edge_index(0, i + 1)];
  for (int i = 0; i < d; ++i)
    for (int j = i + 1; j < d; ++j) {
      double s0i = s[edge_index(0, i + 1)];
      double s0j = s[edge_index(0, j + 1)];
      double sij = s[edge_index(i + 1, j + 1)];
#if ORC_MUTATION == 3
      double val = 0.5 * (s0i + s0j + sij);
#else
      double val = 0.5 * (s0i + s0j - sij);
#endif
      g[i * d + j] = val;
      g[j * d + i] = val;
    }
  double dt = leibniz_det(g, d);
  *det = dt;
  if (!ginv) return;
  if (d == 1) {
    ginv[0] = 1.0 / g[0];
    return;
  }
  double minor[MAXD * MAXD];
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      /* cofactor C_ij = (-1)^(i+j) det(g without row i, col j); inverse = C^T/det */
      int r = 0;
      for (int a = 0; a < d; ++a) {
        if (a == i) continue;
        int c = 0;
        for (int b = 0; b < d; ++b) {
          if (b == j) continue;
          minor[r * (d - 1) + c] = g[a * d + b];
          ++c;
        }
        ++r;
      }
      double cof = leibniz_det(minor, d - 1);
      if ((i + j) & 1) cof = -cof;
      ginv[j * d + i] = cof / dt;
    }
}

/* Shape of WhitneyPairing element matrices (formoniq/src/operators.rs:197-211):
 * the pairing's grade is that of the inner product; a differentiated side sits
 * one grade below. */
void orc_form_shape(int d, int kind, int grade, int *rows, int *cols) {
  int nk = (int)orc_binom(d + 1, grade + 1);
  int nkm = (grade >= 1) ? (int)orc_binom(d + 1, grade) : 0;
  int test_dif = (kind == ORC_DIF_TEST || kind == ORC_DIF_BOTH);
  int trial_dif = (kind == ORC_DIF_TRIAL || kind == ORC_DIF_BOTH);
  *rows = test_dif ? nkm : nk;
  *cols = trial_dif ? nkm : nk;
}

/* HodgeMass::element (formoniq/src/operators.rs:84-94):
 *   M = vol * C^T (H (x) Q) C,  H = P (Lambda^k g^-1) P^T,  P = Lambda^k(Dlambda).
 * The pullback follows WhitneyExpansion::{column, pullback}
 * (derham/src/interpolate/form.rs:195-200, 222-232): entry (i,j) is the sum over
 * the deletion pairs of dof_i and dof_j of a*b*H[blade_a, blade_b]*Q[vert_a, vert_b],
 * with a = (-1)^p k! for the deletion at position p. */
static void hodge_mass(int d, int k, const double *s, double *M) {
  double g[MAXD * MAXD], ginv[MAXD * MAXD], det;
  orc_cell_metric(d, s, g, ginv, &det);

  /* Lambda^k g^-1 over colex k-subsets of {0..d-1} (metric/src/tensor.rs:127-131). */
  unsigned bl[32];
  int nbl = subset_masks(d, k, bl);
  double L[36 * 36];
  double minor[MAXD * MAXD];
  for (int I = 0; I < nbl; ++I)
    for (int J = 0; J < nbl; ++J) {
      for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b)
          minor[a * k + b] = ginv[mask_elem(bl[I], a) * d + mask_elem(bl[J], b)];
      L[I * nbl + J] = leibniz_det(minor, k);
    }

  /* P = exterior_power(unit_difbarys(d), k) (multialgebra/src/lib.rs:493-514;
   * simplicial/src/atlas.rs:174-182): Dlambda row 0 = -1, rows 1..d = I. */
  unsigned pb[32];
  int npb = subset_masks(d + 1, k, pb);
  double P[36 * 36];
  for (int A = 0; A < npb; ++A)
    for (int I = 0; I < nbl; ++I) {
      for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b) {
          int row = mask_elem(pb[A], a), col = mask_elem(bl[I], b);
          minor[a * k + b] = (row == 0) ? -1.0 : ((row - 1 == col) ? 1.0 : 0.0);
        }
      P[A * nbl + I] = leibniz_det(minor, k);
    }
  /* H = (P L) P^T, dense products in natural order. */
  double PL[36 * 36], H[36 * 36];
  for (int A = 0; A < npb; ++A)
    for (int J = 0; J < nbl; ++J) {
      double acc = 0.0;
      for (int I = 0; I < nbl; ++I) acc += P[A * nbl + I] * L[I * nbl + J];
      PL[A * nbl + J] = acc;
    }
  for (int A = 0; A < npb; ++A)
    for (int B = 0; B < npb; ++B) {
      double acc = 0.0;
      for (int J = 0; J < nbl; ++J) acc += PL[A * nbl + J] * P[B * nbl + J];
      H[A * npb + B] = acc;
    }

  /* Q = (1 + delta)/((n+1)(n+2)) at unit volume (simplicial/src/atlas.rs:198-204). */
  int nv = d + 1;
  double qscale = 1.0 / (double)(nv * (nv + 1));
  double Q[MAXV * MAXV];
  for (int v = 0; v < nv; ++v)
    for (int w = 0; w < nv; ++w) Q[v * nv + w] = (v == w && ORC_MUTATION != 1) ? 2.0 * qscale : qscale;

  unsigned dofs[32];
  int ndof = subset_masks(d + 1, k + 1, dofs);
  double kf = factorial(k);
  /* vol = sqrt|det g| / d! (regge/src/lib.rs:26-28, simplicial/src/atlas.rs:102-105) */
#if ORC_MUTATION == 5
  double vol = sqrt(fabs(det));
#else
  double vol = (1.0 / factorial(d)) * sqrt(fabs(det));
#endif
  for (int i = 0; i < ndof; ++i)
    for (int j = 0; j < ndof; ++j) {
      double sum = 0.0;
      for (int p = 0; p <= k; ++p) {
        int va = mask_elem(dofs[i], p);
        unsigned ba = dofs[i] & ~(1u << va);
        double aval = ((p & 1) && ORC_MUTATION != 2 ? -1.0 : 1.0) * kf;
        for (int q = 0; q <= k; ++q) {
          int vb = mask_elem(dofs[j], q);
          unsigned bb = dofs[j] & ~(1u << vb);
          double bval = ((q & 1) ? -1.0 : 1.0) * kf;
          sum += aval * bval * H[mask_rank(ba) * npb + mask_rank(bb)] * Q[va * nv + vb];
        }
      }
      M[i * ndof + j] = vol * sum;
    }
}

/* unit_boundary_operator(d, k) (simplicial/src/topology/simplex.rs:223-238):
 * C(d+1,k) x C(d+1,k+1), entry (rank(face), coface) = (-1)^p for the deletion
 * at position p. Zero rows at k = 0. */
static void unit_boundary(int d, int k, double *B, int *nrows, int *ncols) {
  unsigned co[32];
  int nco = subset_masks(d + 1, k + 1, co);
  int nf = (k >= 1) ? (int)orc_binom(d + 1, k) : 0;
  *nrows = nf;
  *ncols = nco;
  for (int i = 0; i < nf * nco; ++i) B[i] = 0.0;
  if (k < 1) return;
  for (int c = 0; c < nco; ++c)
    for (int p = 0; p <= k; ++p) {
      int v = mask_elem(co[c], p);
      unsigned face = co[c] & ~(1u << v);
      B[mask_rank(face) * nco + c] = ((p & 1) != (ORC_MUTATION == 4)) ? -1.0 : 1.0;
    }
}

/* WhitneyPairing::element (formoniq/src/operators.rs:201-211):
 *   mass, mass * D (dif_trial), dbd * mass (dif_test), dbd * mass * D (dif_both),
 * with D = boundary^T; nalgebra dense products in natural order. */
void orc_element_matrix(int d, int kind, int grade, const double *s, double *out) {
  double M[10 * 10];
  hodge_mass(d, grade, s, M);
  int n = (int)orc_binom(d + 1, grade + 1);
  double Bd[10 * 10];
  int br = 0, bc = 0;
  unit_boundary(d, grade, Bd, &br, &bc);
  double MD[10 * 10];
  int rows = n, cols = n;
  const double *cur = M;
  if (kind == ORC_DIF_TRIAL || kind == ORC_DIF_BOTH) {
    /* (M * B^T): n x br */
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < br; ++j) {
        double acc = 0.0;
        for (int l = 0; l < n; ++l) acc += M[i * n + l] * Bd[j * bc + l];
        MD[i * br + j] = acc;
      }
    cur = MD;
    cols = br;
  }
  if (kind == ORC_DIF_TEST || kind == ORC_DIF_BOTH) {
    for (int i = 0; i < br; ++i)
      for (int j = 0; j < cols; ++j) {
        double acc = 0.0;
        for (int l = 0; l < n; ++l) acc += Bd[i * bc + l] * cur[l * cols + j];
        out[i * cols + j] = acc;
      }
    rows = br;
    (void)rows;
    return;
  }
  for (int i = 0; i < rows * cols; ++i) out[i] = cur[i];
}

/* ------------------------------------------------------------------------ */
/* meshers                                                                   */
/* ------------------------------------------------------------------------ */

int64_t orc_kuhn_ncells(int d, int n) {
  int64_t c = 1;
  for (int i = 0; i < d; ++i) c *= n;
  return c * (int64_t)factorial(d);
}
int64_t orc_kuhn_nverts(int d, int n) {
  int64_t c = 1;
  for (int i = 0; i < d; ++i) c *= (n + 1);
  return c;
}

/* CartesianTopology::cell_skeleton (simplicial/src/mesher/grid.rs:77-98): for
 * each box (axis 0 fastest) and each axis permutation, the chain origin,
 * origin + e_pi0, ...; vertex ids axis-0-fastest (multiindex/src/cartesian.rs:94-105).
 * CartesianGrid::vertex_pos (regge/src/mesher/cartesian.rs:158-168): c_a / n. */
void orc_kuhn(int d, int n, int64_t *cells, double *coords) {
  int64_t nv = orc_kuhn_nverts(d, n);
  int64_t stride[MAXD];
  int64_t s = 1;
  for (int a = 0; a < d; ++a) {
    stride[a] = s;
    s *= (n + 1);
  }
  if (coords)
    for (int64_t v = 0; v < nv; ++v) {
      int64_t r = v;
      for (int a = 0; a < d; ++a) {
        int64_t c = r % (n + 1);
        r /= (n + 1);
        coords[v * d + a] = (double)c / (double)n;
      }
    }
  if (!cells) return;
  int64_t nbox = 1;
  for (int a = 0; a < d; ++a) nbox *= n;
  int perm[MAXD];
  int64_t out = 0;
  for (int64_t b = 0; b < nbox; ++b) {
    int64_t r = b, origin = 0;
    for (int a = 0; a < d; ++a) {
      origin += (r % n) * stride[a];
      r /= n;
    }
    for (int i = 0; i < d; ++i) perm[i] = i;
    for (;;) {
      int64_t v = origin;
      cells[out * (d + 1)] = v;
      for (int i = 0; i < d; ++i) {
        v += stride[perm[i]];
        cells[out * (d + 1) + i + 1] = v;
      }
      ++out;
      int piv = -1;
      for (int i = d - 2; i >= 0; --i)
        if (perm[i] < perm[i + 1]) {
          piv = i;
          break;
        }
      if (piv < 0) break;
      int succ = d - 1;
      while (perm[succ] <= perm[piv]) --succ;
      int t = perm[piv];
      perm[piv] = perm[succ];
      perm[succ] = t;
      for (int i = piv + 1, j = d - 1; i < j; ++i, --j) {
        t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* topology                                                                  */
/* ------------------------------------------------------------------------ */

struct orc_complex {
  int d;
  int64_t n[MAXV];        /* simplices per grade */
  int64_t *simp[MAXV];    /* [n_k][k+1], colex sorted */
  int64_t *faces[MAXV];   /* [ncells][C(d+1,k+1)] global kidx of local faces */
  uint8_t *bnd[MAXV];     /* boundary flags per grade */
};

typedef struct {
  unsigned __int128 key;
  int64_t slot;
} keyslot;

static int cmp_keyslot(const void *a, const void *b) {
  const keyslot *x = (const keyslot *)a, *y = (const keyslot *)b;
  if (x->key < y->key) return -1;
  if (x->key > y->key) return 1;
  return (x->slot < y->slot) ? -1 : (x->slot > y->slot);
}

/* Key whose integer order is the colex order of the simplex: the reversed
 * vertex tuple read in base V (simplicial/src/topology/simplex.rs:126-130). */
static unsigned __int128 colex_key(const int64_t *v, int len, int64_t V) {
  unsigned __int128 key = 0;
  for (int i = len - 1; i >= 0; --i) key = key * (unsigned __int128)V + (unsigned __int128)v[i];
  return key;
}

static int cmp_i64(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}

/* Complex::from_cells_unchecked (simplicial/src/topology/complex.rs:298-336)
 * with Skeleton::new (skeleton.rs:50-83): every k-skeleton is the colex-sorted,
 * deduplicated set of (k+1)-subsets of all cells; the cells are sorted too. */
orc_complex *orc_complex_new(int d, int64_t ncells, const int64_t *cells) {
  if (d < 1 || d > MAXD || ncells <= 0) return NULL;
  int nv = d + 1;
  int64_t *cs = (int64_t *)malloc(sizeof(int64_t) * ncells * nv);
  memcpy(cs, cells, sizeof(int64_t) * ncells * nv);
  int64_t V = 0;
  for (int64_t c = 0; c < ncells; ++c) {
    qsort(cs + c * nv, nv, sizeof(int64_t), cmp_i64);
    for (int i = 0; i < nv; ++i) {
      if (cs[c * nv + i] < 0 || (i > 0 && cs[c * nv + i] == cs[c * nv + i - 1])) {
        free(cs);
        return NULL;
      }
      if (cs[c * nv + i] + 1 > V) V = cs[c * nv + i] + 1;
    }
  }
  /* vertices contiguous and used */
  uint8_t *used = (uint8_t *)calloc(V, 1);
  for (int64_t i = 0; i < ncells * nv; ++i) used[cs[i]] = 1;
  for (int64_t v = 0; v < V; ++v)
    if (!used[v]) {
      free(used);
      free(cs);
      return NULL;
    }
  free(used);

  orc_complex *cx = (orc_complex *)calloc(1, sizeof(orc_complex));
  cx->d = d;
  /* sort cells colex, dedup */
  keyslot *ks = (keyslot *)malloc(sizeof(keyslot) * ncells);
  for (int64_t c = 0; c < ncells; ++c) {
    ks[c].key = colex_key(cs + c * nv, nv, V);
    ks[c].slot = c;
  }
  qsort(ks, ncells, sizeof(keyslot), cmp_keyslot);
  int64_t nc = 0;
  int64_t *cells_sorted = (int64_t *)malloc(sizeof(int64_t) * ncells * nv);
  for (int64_t i = 0; i < ncells; ++i) {
    if (i > 0 && ks[i].key == ks[i - 1].key) continue;
    memcpy(cells_sorted + nc * nv, cs + ks[i].slot * nv, sizeof(int64_t) * nv);
    ++nc;
  }
  free(ks);
  free(cs);
  cx->n[d] = nc;
  cx->simp[d] = cells_sorted;
  cx->faces[d] = (int64_t *)malloc(sizeof(int64_t) * nc);
  for (int64_t c = 0; c < nc; ++c) cx->faces[d][c] = c;

  for (int k = 0; k < d; ++k) {
    unsigned masks[32];
    int nf = subset_masks(nv, k + 1, masks);
    int64_t total = nc * nf;
    keyslot *e = (keyslot *)malloc(sizeof(keyslot) * total);
    int64_t tmp[MAXV];
    for (int64_t c = 0; c < nc; ++c)
      for (int f = 0; f < nf; ++f) {
        for (int p = 0; p <= k; ++p) tmp[p] = cells_sorted[c * nv + mask_elem(masks[f], p)];
        e[c * nf + f].key = colex_key(tmp, k + 1, V);
        e[c * nf + f].slot = c * nf + f;
      }
    qsort(e, total, sizeof(keyslot), cmp_keyslot);
    int64_t *fc = (int64_t *)malloc(sizeof(int64_t) * total);
    int64_t cnt = 0;
    for (int64_t i = 0; i < total; ++i) {
      if (i == 0 || e[i].key != e[i - 1].key) ++cnt;
      fc[e[i].slot] = cnt - 1;
    }
    int64_t *sk = (int64_t *)malloc(sizeof(int64_t) * cnt * (k + 1));
    for (int64_t i = 0; i < total; ++i) {
      int64_t c = e[i].slot / nf;
      int f = (int)(e[i].slot % nf);
      int64_t kidx = fc[e[i].slot];
      for (int p = 0; p <= k; ++p) sk[kidx * (k + 1) + p] = cells_sorted[c * nv + mask_elem(masks[f], p)];
    }
    free(e);
    cx->n[k] = cnt;
    cx->simp[k] = sk;
    cx->faces[k] = fc;
  }
  if (cx->n[0] != V) { /* 0-skeleton must be the vertices 0..V-1 (skeleton.rs:63-68) */
    orc_complex_free(cx);
    return NULL;
  }

  /* Boundary: facets in exactly one cell, and every face of those
   * (simplicial/src/topology/complex.rs:140-149). */
  int nfac = nv; /* C(d+1, d) */
  unsigned fmask[32];
  subset_masks(nv, d, fmask);
  int64_t *mult = (int64_t *)calloc(cx->n[d - 1], sizeof(int64_t));
  for (int64_t i = 0; i < nc * nfac; ++i) mult[cx->faces[d - 1][i]]++;
  for (int k = 0; k <= d; ++k) cx->bnd[k] = (uint8_t *)calloc(cx->n[k] ? cx->n[k] : 1, 1);
  for (int64_t c = 0; c < nc; ++c)
    for (int f = 0; f < nfac; ++f) {
      if (mult[cx->faces[d - 1][c * nfac + f]] != 1) continue;
      for (int k = 0; k < d; ++k) {
        unsigned sm[32];
        int ns = subset_masks(nv, k + 1, sm);
        for (int s2 = 0; s2 < ns; ++s2)
          if ((sm[s2] & ~fmask[f]) == 0) cx->bnd[k][cx->faces[k][c * ns + s2]] = 1;
      }
    }
  free(mult);
  return cx;
}

void orc_complex_free(orc_complex *c) {
  if (!c) return;
  for (int k = 0; k < MAXV; ++k) {
    free(c->simp[k]);
    free(c->faces[k]);
    free(c->bnd[k]);
  }
  free(c);
}

int orc_dim(const orc_complex *c) { return c->d; }
int64_t orc_nsimplices(const orc_complex *c, int k) {
  return (k < 0 || k > c->d) ? 0 : c->n[k];
}
void orc_skeleton(const orc_complex *c, int k, int64_t *out) {
  memcpy(out, c->simp[k], sizeof(int64_t) * c->n[k] * (k + 1));
}
void orc_cell_faces(const orc_complex *c, int k, int64_t *out) {
  int nf = (int)orc_binom(c->d + 1, k + 1);
  memcpy(out, c->faces[k], sizeof(int64_t) * c->n[c->d] * nf);
}
void orc_boundary_flags(const orc_complex *c, int k, uint8_t *out) {
  memcpy(out, c->bnd[k], c->n[k]);
}

/* ------------------------------------------------------------------------ */
/* CSR (nalgebra-sparse semantics)                                           */
/* ------------------------------------------------------------------------ */

orc_csr *orc_csr_new(int64_t nrows, int64_t ncols, int64_t nnz) {
  orc_csr *a = (orc_csr *)calloc(1, sizeof(orc_csr));
  a->nrows = nrows;
  a->ncols = ncols;
  a->nnz = nnz;
  a->row_ptr = (int64_t *)calloc(nrows + 1, sizeof(int64_t));
  a->col = (int64_t *)malloc(sizeof(int64_t) * (nnz ? nnz : 1));
  a->val = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
  return a;
}
void orc_csr_free(orc_csr *a) {
  if (!a) return;
  free(a->row_ptr);
  free(a->col);
  free(a->val);
  free(a);
}

/* COO -> CSR: rows ascending, columns ascending within a row, duplicates summed
 * in insertion order, summed zeros kept (CsrMatrix::from(&coo),
 * formoniq/src/galerkin.rs:184-187). */
static orc_csr *coo_to_csr(int64_t nrows, int64_t ncols, int64_t n, const int64_t *r,
                           const int64_t *c, const double *v) {
  int64_t *cnt = (int64_t *)calloc(nrows + 1, sizeof(int64_t));
  for (int64_t i = 0; i < n; ++i) cnt[r[i] + 1]++;
  for (int64_t i = 0; i < nrows; ++i) cnt[i + 1] += cnt[i];
  int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (nrows + 1));
  memcpy(pos, cnt, sizeof(int64_t) * (nrows + 1));
  int64_t *oc = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
  double *ov = (double *)malloc(sizeof(double) * (n ? n : 1));
  for (int64_t i = 0; i < n; ++i) {
    int64_t p = pos[r[i]]++;
    oc[p] = c[i];
    ov[p] = v[i];
  }
  free(pos);
  /* stable insertion sort by column inside each row, then combine */
  int64_t nnz = 0;
  int64_t *rp = (int64_t *)calloc(nrows + 1, sizeof(int64_t));
  for (int64_t row = 0; row < nrows; ++row) {
    int64_t b = cnt[row], e = cnt[row + 1];
    for (int64_t i = b + 1; i < e; ++i) {
      int64_t kc = oc[i];
      double kv = ov[i];
      int64_t j = i - 1;
      while (j >= b && oc[j] > kc) {
        oc[j + 1] = oc[j];
        ov[j + 1] = ov[j];
        --j;
      }
      oc[j + 1] = kc;
      ov[j + 1] = kv;
    }
    for (int64_t i = b; i < e;) {
      int64_t col = oc[i];
      double sum = ov[i];
      int64_t j = i + 1;
      while (j < e && oc[j] == col) sum += ov[j++];
      oc[nnz] = col;
      ov[nnz] = sum;
      ++nnz;
      i = j;
    }
    rp[row + 1] = nnz;
  }
  free(cnt);
  orc_csr *a = (orc_csr *)calloc(1, sizeof(orc_csr));
  a->nrows = nrows;
  a->ncols = ncols;
  a->nnz = nnz;
  a->row_ptr = rp;
  a->col = oc;
  a->val = ov;
  return a;
}

/* Sparse addition: union pattern, a + b (hdif_gram, whitney_complex.rs:180-183). */
orc_csr *orc_csr_add(const orc_csr *a, const orc_csr *b) {
  orc_csr *o = orc_csr_new(a->nrows, a->ncols, a->nnz + b->nnz);
  int64_t nnz = 0;
  for (int64_t r = 0; r < a->nrows; ++r) {
    int64_t i = a->row_ptr[r], ie = a->row_ptr[r + 1];
    int64_t j = b->row_ptr[r], je = b->row_ptr[r + 1];
    while (i < ie || j < je) {
      if (j >= je || (i < ie && a->col[i] < b->col[j])) {
        o->col[nnz] = a->col[i];
        o->val[nnz++] = a->val[i++];
      } else if (i >= ie || b->col[j] < a->col[i]) {
        o->col[nnz] = b->col[j];
        o->val[nnz++] = b->val[j++];
      } else {
        o->col[nnz] = a->col[i];
        o->val[nnz++] = a->val[i++] + b->val[j++];
      }
    }
    o->row_ptr[r + 1] = nnz;
  }
  o->nnz = nnz;
  return o;
}

/* Structural CSR x CSR: for each row i, for each (i,k,a) ascending k, for each
 * (k,j,b): c_ij += a*b; every touched (i,j) kept (dirac.rs:244). */
orc_csr *orc_csr_mul(const orc_csr *a, const orc_csr *b) {
  int64_t *mark = (int64_t *)malloc(sizeof(int64_t) * (b->ncols ? b->ncols : 1));
  double *acc = (double *)calloc(b->ncols ? b->ncols : 1, sizeof(double));
  for (int64_t j = 0; j < b->ncols; ++j) mark[j] = -1;
  int64_t cap = a->nnz * 4 + 16, nnz = 0;
  int64_t *col = (int64_t *)malloc(sizeof(int64_t) * cap);
  double *val = (double *)malloc(sizeof(double) * cap);
  int64_t *rp = (int64_t *)calloc(a->nrows + 1, sizeof(int64_t));
  int64_t *list = (int64_t *)malloc(sizeof(int64_t) * (b->ncols ? b->ncols : 1));
  for (int64_t i = 0; i < a->nrows; ++i) {
    int64_t nl = 0;
    for (int64_t p = a->row_ptr[i]; p < a->row_ptr[i + 1]; ++p) {
      int64_t k = a->col[p];
      double av = a->val[p];
      for (int64_t q = b->row_ptr[k]; q < b->row_ptr[k + 1]; ++q) {
        int64_t j = b->col[q];
        if (mark[j] != i) {
          mark[j] = i;
          acc[j] = 0.0;
          list[nl++] = j;
        }
        acc[j] += av * b->val[q];
      }
    }
    qsort(list, nl, sizeof(int64_t), cmp_i64);
    if (nnz + nl > cap) {
      while (nnz + nl > cap) cap *= 2;
      col = (int64_t *)realloc(col, sizeof(int64_t) * cap);
      val = (double *)realloc(val, sizeof(double) * cap);
    }
    for (int64_t t = 0; t < nl; ++t) {
      col[nnz] = list[t];
      val[nnz++] = acc[list[t]];
    }
    rp[i + 1] = nnz;
  }
  free(mark);
  free(acc);
  free(list);
  orc_csr *o = (orc_csr *)calloc(1, sizeof(orc_csr));
  o->nrows = a->nrows;
  o->ncols = b->ncols;
  o->nnz = nnz;
  o->row_ptr = rp;
  o->col = col;
  o->val = val;
  return o;
}

orc_csr *orc_csr_transpose(const orc_csr *a) {
  orc_csr *t = orc_csr_new(a->ncols, a->nrows, a->nnz);
  for (int64_t p = 0; p < a->nnz; ++p) t->row_ptr[a->col[p] + 1]++;
  for (int64_t i = 0; i < a->ncols; ++i) t->row_ptr[i + 1] += t->row_ptr[i];
  int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (a->ncols + 1));
  memcpy(pos, t->row_ptr, sizeof(int64_t) * (a->ncols + 1));
  for (int64_t r = 0; r < a->nrows; ++r)
    for (int64_t p = a->row_ptr[r]; p < a->row_ptr[r + 1]; ++p) {
      int64_t q = pos[a->col[p]]++;
      t->col[q] = r;
      t->val[q] = a->val[p];
    }
  free(pos);
  return t;
}

/* E^T A E for 0/1 selections (RelativeWhitneyComplex::assemble,
 * formoniq/src/whitney_complex.rs:620-624): the kept rows and columns,
 * renumbered in kidx order. */
orc_csr *orc_csr_select(const orc_csr *a, const uint8_t *row_keep, const uint8_t *col_keep) {
  int64_t *cmap = (int64_t *)malloc(sizeof(int64_t) * (a->ncols ? a->ncols : 1));
  int64_t nc = 0, nr = 0;
  for (int64_t j = 0; j < a->ncols; ++j) cmap[j] = (!col_keep || col_keep[j]) ? nc++ : -1;
  for (int64_t i = 0; i < a->nrows; ++i)
    if (!row_keep || row_keep[i]) ++nr;
  orc_csr *o = orc_csr_new(nr, nc, a->nnz);
  int64_t nnz = 0, r = 0;
  for (int64_t i = 0; i < a->nrows; ++i) {
    if (row_keep && !row_keep[i]) continue;
    for (int64_t p = a->row_ptr[i]; p < a->row_ptr[i + 1]; ++p)
      if (cmap[a->col[p]] >= 0) {
        o->col[nnz] = cmap[a->col[p]];
        o->val[nnz++] = a->val[p];
      }
    o->row_ptr[++r] = nnz;
  }
  o->nnz = nnz;
  free(cmap);
  return o;
}

orc_csr *orc_csr_scale_rows(const orc_csr *a, const double *s) {
  orc_csr *o = orc_csr_new(a->nrows, a->ncols, a->nnz);
  memcpy(o->row_ptr, a->row_ptr, sizeof(int64_t) * (a->nrows + 1));
  memcpy(o->col, a->col, sizeof(int64_t) * a->nnz);
  for (int64_t i = 0; i < a->nrows; ++i)
    for (int64_t p = a->row_ptr[i]; p < a->row_ptr[i + 1]; ++p) o->val[p] = s[i] * a->val[p];
  return o;
}

/* CooMatrix::block (used by HodgeBlocks::mixed_hodge_laplacian,
 * formoniq/src/hodge.rs:93-99): blocks placed at offsets, then compressed. */
orc_csr *orc_csr_block(int nb, const orc_csr *const *blocks, const double *scales,
                       const int64_t *row_sizes, const int64_t *col_sizes) {
  int64_t nrows = 0, ncols = 0, nnz = 0;
  int64_t roff[16], coff[16];
  for (int i = 0; i < nb; ++i) {
    roff[i] = nrows;
    nrows += row_sizes[i];
    coff[i] = ncols;
    ncols += col_sizes[i];
  }
  for (int i = 0; i < nb * nb; ++i)
    if (blocks[i]) nnz += blocks[i]->nnz;
  int64_t *r = (int64_t *)malloc(sizeof(int64_t) * (nnz ? nnz : 1));
  int64_t *c = (int64_t *)malloc(sizeof(int64_t) * (nnz ? nnz : 1));
  double *v = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
  int64_t t = 0;
  for (int bi = 0; bi < nb; ++bi)
    for (int bj = 0; bj < nb; ++bj) {
      const orc_csr *b = blocks[bi * nb + bj];
      if (!b) continue;
      double s = scales ? scales[bi * nb + bj] : 1.0;
      for (int64_t i = 0; i < b->nrows; ++i)
        for (int64_t p = b->row_ptr[i]; p < b->row_ptr[i + 1]; ++p) {
          r[t] = roff[bi] + i;
          c[t] = coff[bj] + b->col[p];
          v[t] = (s == 1.0) ? b->val[p] : s * b->val[p];
          ++t;
        }
    }
  orc_csr *o = coo_to_csr(nrows, ncols, t, r, c, v);
  free(r);
  free(c);
  free(v);
  return o;
}

void orc_spmv(const orc_csr *a, const double *x, double *y) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (a->nnz > 200000)
#endif
  for (int64_t i = 0; i < a->nrows; ++i) {
    double acc = 0.0;
    for (int64_t p = a->row_ptr[i]; p < a->row_ptr[i + 1]; ++p) acc += a->val[p] * x[a->col[p]];
    y[i] = acc;
  }
}

/* ------------------------------------------------------------------------ */
/* incidence, orientation, lengths                                           */
/* ------------------------------------------------------------------------ */

/* coboundary_operator(k) = boundary_operator(k+1)^T, with
 * boundary[face, coface] = (-1)^p for the deletion at position p
 * (simplicial/src/topology/complex.rs:168-217). Rows: (k+1)-simplices. */
orc_csr *orc_coboundary(const orc_complex *cx, int k) {
  int d = cx->d;
  if (k < 0 || k >= d) return orc_csr_new(k + 1 <= d ? cx->n[k + 1] : 0, k >= 0 && k <= d ? cx->n[k] : 0, 0);
  int64_t nco = cx->n[k + 1];
  int64_t nnz = nco * (k + 2);
  int64_t *r = (int64_t *)malloc(sizeof(int64_t) * nnz);
  int64_t *c = (int64_t *)malloc(sizeof(int64_t) * nnz);
  double *v = (double *)malloc(sizeof(double) * nnz);
  uint8_t *seen = (uint8_t *)calloc(nco, 1);
  unsigned co[32], fa[32];
  int nv = d + 1;
  int nco_l = subset_masks(nv, k + 2, co);
  int nfa_l = subset_masks(nv, k + 1, fa);
  int64_t t = 0;
  for (int64_t cell = 0; cell < cx->n[d]; ++cell)
    for (int s = 0; s < nco_l; ++s) {
      int64_t gco = cx->faces[k + 1][cell * nco_l + s];
      if (seen[gco]) continue;
      seen[gco] = 1;
      for (int p = 0; p <= k + 1; ++p) {
        unsigned face = co[s] & ~(1u << mask_elem(co[s], p));
        int lf = 0;
        for (int q = 0; q < nfa_l; ++q)
          if (fa[q] == face) lf = q;
        r[t] = gco;
        c[t] = cx->faces[k][cell * nfa_l + lf];
        v[t] = (p & 1) ? -1.0 : 1.0;
        ++t;
      }
    }
  free(seen);
  orc_csr *o = coo_to_csr(nco, cx->n[k], t, r, c, v);
  free(r);
  free(c);
  free(v);
  return o;
}

/* Coherent orientation by propagation across interior facets: neighbours
 * induce opposite orientations on a shared facet. */
int orc_orientation(const orc_complex *cx, int8_t *signs) {
  int d = cx->d, nv = d + 1;
  int64_t nc = cx->n[d], nf = cx->n[d - 1];
  int64_t *a = (int64_t *)malloc(sizeof(int64_t) * nf * 2);
  int8_t *ap = (int8_t *)malloc(nf * 2);
  for (int64_t i = 0; i < nf * 2; ++i) a[i] = -1;
  unsigned fm[32];
  subset_masks(nv, d, fm);
  for (int64_t c = 0; c < nc; ++c)
    for (int f = 0; f < nv; ++f) {
      int64_t g = cx->faces[d - 1][c * nv + f];
      int p = 0; /* deleted position of local facet f */
      for (int q = 0; q < nv; ++q)
        if (!(fm[f] & (1u << q))) p = q;
      int slot = (a[g * 2] < 0) ? 0 : 1;
      a[g * 2 + slot] = c;
      ap[g * 2 + slot] = (int8_t)((p & 1) ? -1 : 1);
    }
  for (int64_t c = 0; c < nc; ++c) signs[c] = 0;
  int64_t *stack = (int64_t *)malloc(sizeof(int64_t) * nc);
  int ok = 1;
  for (int64_t root = 0; root < nc && ok; ++root) {
    if (signs[root]) continue;
    int64_t top = 0;
    signs[root] = 1;
    stack[top++] = root;
    while (top && ok) {
      int64_t c = stack[--top];
      for (int f = 0; f < nv; ++f) {
        int64_t g = cx->faces[d - 1][c * nv + f];
        if (a[g * 2 + 1] < 0) continue;
        int me = (a[g * 2] == c) ? 0 : 1;
        int64_t o = a[g * 2 + 1 - me];
        int8_t want = (int8_t)(-signs[c] * ap[g * 2 + me] * ap[g * 2 + 1 - me]);
        if (!signs[o]) {
          signs[o] = want;
          stack[top++] = o;
        } else if (signs[o] != want) {
          ok = 0;
        }
      }
    }
  }
  free(stack);
  free(a);
  free(ap);
  return ok;
}

/* MeshCoords::to_edge_lengths_sq (regge/src/coord/mesh.rs:208-216): the ambient
 * squared norm of v_j - v_i per edge in kidx order; Minkowski is diag(-1,1,..). */
void orc_lengths_from_coords(const orc_complex *cx, const double *coords, int minkowski,
                             double *out) {
  int d = cx->d;
  for (int64_t e = 0; e < cx->n[1]; ++e) {
    int64_t vi = cx->simp[1][2 * e], vj = cx->simp[1][2 * e + 1];
    double acc = 0.0;
    for (int a = 0; a < d; ++a) {
      double t = coords[vj * d + a] - coords[vi * d + a];
      acc += (minkowski && a == 0) ? -(t * t) : t * t;
    }
    out[e] = acc;
  }
}

/* ------------------------------------------------------------------------ */
/* element matrices and assembly                                             */
/* ------------------------------------------------------------------------ */

static void gather_lengths(const orc_complex *cx, const double *lengths, int64_t c, double *s) {
  int ne = (int)orc_binom(cx->d + 1, 2);
  for (int e = 0; e < ne; ++e) s[e] = lengths[cx->faces[1][c * ne + e]];
}

void orc_element_matrices(const orc_complex *cx, const double *lengths, int kind, int grade,
                          double *out, int nthreads) {
  int rows, cols;
  orc_form_shape(cx->d, kind, grade, &rows, &cols);
  int64_t nc = cx->n[cx->d];
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
  for (int64_t c = 0; c < nc; ++c) {
    double s[10];
    gather_lengths(cx, lengths, c, s);
    orc_element_matrix(cx->d, kind, grade, s, out + c * rows * cols);
  }
  (void)nthreads;
}

/* assemble_matrix (formoniq/src/galerkin.rs:138-188): per cell, metric ->
 * element -> push (kidx_i, kidx_j, v) when v != 0.0, triplets concatenated in
 * cell order; then COO -> CSR. */
orc_csr *orc_assemble(const orc_complex *cx, const double *lengths, int kind, int grade,
                      int nthreads) {
  int d = cx->d;
  int rows, cols;
  orc_form_shape(d, kind, grade, &rows, &cols);
  int tg = (kind == ORC_DIF_TEST || kind == ORC_DIF_BOTH) ? grade - 1 : grade;
  int rg = (kind == ORC_DIF_TRIAL || kind == ORC_DIF_BOTH) ? grade - 1 : grade;
  int64_t nr = orc_nsimplices(cx, tg), ncl = orc_nsimplices(cx, rg);
  if (nr == 0 || ncl == 0 || rows == 0 || cols == 0) return orc_csr_new(nr, ncl, 0);
  int64_t nc = cx->n[d];
  int per = rows * cols;
  double *el = (double *)malloc(sizeof(double) * nc * per);
  orc_element_matrices(cx, lengths, kind, grade, el, nthreads);
  int64_t *r = (int64_t *)malloc(sizeof(int64_t) * nc * per);
  int64_t *c = (int64_t *)malloc(sizeof(int64_t) * nc * per);
  double *v = (double *)malloc(sizeof(double) * nc * per);
  int nft = (int)orc_binom(d + 1, tg + 1), nfr = (int)orc_binom(d + 1, rg + 1);
  int64_t t = 0;
  for (int64_t cell = 0; cell < nc; ++cell)
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < cols; ++j) {
        double val = el[cell * per + i * cols + j];
        if (val != 0.0) {
          r[t] = cx->faces[tg][cell * nft + i];
          c[t] = cx->faces[rg][cell * nfr + j];
          v[t] = val;
          ++t;
        }
      }
  free(el);
  orc_csr *o = coo_to_csr(nr, ncl, t, r, c, v);
  free(r);
  free(c);
  free(v);
  return o;
}

/* ------------------------------------------------------------------------ */
/* Krylov solvers                                                            */
/* ------------------------------------------------------------------------ */

static double dot(const double *a, const double *b, int64_t n) {
  double s = 0.0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(+ : s) if (n > 100000)
#endif
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

static void precond_apply(const double *inv_diag, const double *r, double *z, int64_t n) {
  if (!inv_diag) {
    memcpy(z, r, sizeof(double) * n);
    return;
  }
  for (int64_t i = 0; i < n; ++i) z[i] = r[i] * inv_diag[i];
}

/* Preconditioned CG from x0 = 0 (iterative/src/krylov.rs:48-95): the relative
 * residual |r|/|b| is checked before every step. */
orc_report orc_cg(const orc_csr *a, const double *inv_diag, const double *b, double rtol,
                  int64_t max_iters, double *x) {
  int64_t n = a->nrows;
  orc_report rep = {0, 0.0, 0};
  double bn = sqrt(dot(b, b, n));
  memset(x, 0, sizeof(double) * n);
  if (bn == 0.0) {
    rep.converged = 1;
    return rep;
  }
  double *r = (double *)malloc(sizeof(double) * n);
  double *z = (double *)malloc(sizeof(double) * n);
  double *p = (double *)malloc(sizeof(double) * n);
  double *ap = (double *)malloc(sizeof(double) * n);
  memcpy(r, b, sizeof(double) * n);
  precond_apply(inv_diag, r, z, n);
  memcpy(p, z, sizeof(double) * n);
  double rz = dot(r, z, n);
  int64_t iters = 0;
  double res;
  for (;;) {
    res = sqrt(dot(r, r, n)) / bn;
    if (res <= rtol || iters >= max_iters) break;
    orc_spmv(a, p, ap);
    double alpha = rz / dot(p, ap, n);
    for (int64_t i = 0; i < n; ++i) x[i] += alpha * p[i];
    for (int64_t i = 0; i < n; ++i) r[i] += -alpha * ap[i];
    precond_apply(inv_diag, r, z, n);
    double rzn = dot(r, z, n);
    double beta = rzn / rz;
    for (int64_t i = 0; i < n; ++i) p[i] = beta * p[i] + z[i];
    rz = rzn;
    ++iters;
  }
  rep.iters = iters;
  rep.residual = res;
  rep.converged = res <= rtol;
  free(r);
  free(z);
  free(p);
  free(ap);
  return rep;
}

/* A preconditioner z = B r: Jacobi (diag) or the Chebyshev block polynomial. */
typedef struct precond {
  const double *inv_diag;  /* Jacobi when cheb == NULL */
  const orc_csr *cheb_p;   /* block-diagonal hdif_gram matrix, np x np */
  const double *coef;      /* 2 + 4 (m-1): 1/theta per block, then (a, b) per step and block */
  int64_t split;           /* rows < split belong to block 0 */
  int m;
  double *res, *d, *dn, *t; /* np work vectors */
} precond;

/* Chebyshev iteration (Saad, Iterative Methods for Sparse Linear Systems,
 * Algorithm 12.1) on D^-1 P from z = 0, m directions summed; rows >= np are
 * the identity. */
static void cheb_apply(const precond *pc, const double *r, double *z, int64_t n) {
  const orc_csr *P = pc->cheb_p;
  const int64_t np = P->nrows;
  const double *dinv = pc->inv_diag, *c = pc->coef;
  double *res = pc->res, *d = pc->d, *dn = pc->dn, *t = pc->t;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (np > 100000)
#endif
  for (int64_t i = 0; i < np; ++i) {
    res[i] = dinv[i] * r[i];
    d[i] = res[i] * c[i < pc->split ? 0 : 1];
    z[i] = 0.0;
  }
  for (int k = 0; k + 1 < pc->m; ++k) {
    orc_spmv(P, d, t);
    const double *ck = c + 2 + 4 * k;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (np > 100000)
#endif
    for (int64_t i = 0; i < np; ++i) {
      const int blk = i < pc->split ? 0 : 1;
      res[i] -= dinv[i] * t[i];
      z[i] += d[i];
      dn[i] = ck[2 * blk] * d[i] + ck[2 * blk + 1] * res[i];
    }
    double *sw = d;
    d = dn;
    dn = sw;
  }
  for (int64_t i = 0; i < np; ++i) z[i] += d[i];
  for (int64_t i = np; i < n; ++i) z[i] = r[i];
}

static void precond_run(const precond *pc, const double *r, double *z, int64_t n) {
  if (pc->cheb_p)
    cheb_apply(pc, r, z, n);
  else
    precond_apply(pc->inv_diag, r, z, n);
}

/* Preconditioned MINRES (Paige-Saunders) from x0 = 0 (iterative/src/krylov.rs:113-211);
 * stops on phibar/beta1 <= rtol, the preconditioner-norm relative residual. */
static orc_report minres_core(const orc_csr *a, const precond *pc, const double *b, double rtol,
                              int64_t max_iters, double *x) {
  int64_t n = a->nrows;
  orc_report rep = {0, 1.0, 0};
  double *r1 = (double *)malloc(sizeof(double) * n);
  double *r2 = (double *)malloc(sizeof(double) * n);
  double *y = (double *)malloc(sizeof(double) * n);
  double *v = (double *)malloc(sizeof(double) * n);
  double *w = (double *)calloc(n, sizeof(double));
  double *w2 = (double *)calloc(n, sizeof(double));
  double *yn = (double *)malloc(sizeof(double) * n);
  memcpy(r1, b, sizeof(double) * n);
  precond_run(pc, r1, y, n);
  double beta1_sq = dot(r1, y, n);
  memset(x, 0, sizeof(double) * n);
  if (beta1_sq <= 0.0) {
    rep.residual = 0.0;
    rep.converged = 1;
    goto done;
  }
  {
    double eps = 2.220446049250313e-16;
    double beta1 = sqrt(beta1_sq);
    double oldb = 0.0, beta = beta1, dbar = 0.0, epsln = 0.0, phibar = beta1;
    double cs = -1.0, sn = 0.0;
    memcpy(r2, r1, sizeof(double) * n);
    double residual = 1.0;
    int64_t iters = 0;
    int converged = 0;
    while (iters < max_iters) {
      ++iters;
      double s = 1.0 / beta;
      for (int64_t i = 0; i < n; ++i) v[i] = y[i] * s;
      orc_spmv(a, v, yn);
      if (iters >= 2) {
        double f = -beta / oldb;
        for (int64_t i = 0; i < n; ++i) yn[i] += f * r1[i];
      }
      double alfa = dot(v, yn, n);
      double f2 = -alfa / beta;
      for (int64_t i = 0; i < n; ++i) yn[i] += f2 * r2[i];
      double *t = r1;
      r1 = r2;
      r2 = yn;
      yn = t;
      precond_run(pc, r2, y, n);
      oldb = beta;
      double bb = dot(r2, y, n);
      beta = sqrt(bb > 0.0 ? bb : 0.0);

      double oldeps = epsln;
      double delta = cs * dbar + sn * alfa;
      double gbar = sn * dbar - cs * alfa;
      epsln = sn * beta;
      dbar = -cs * beta;
      double gamma = sqrt(gbar * gbar + beta * beta);
      if (gamma < eps) gamma = eps;
      cs = gbar / gamma;
      sn = beta / gamma;
      double phi = cs * phibar;
      phibar *= sn;

      double ig = 1.0 / gamma;
      for (int64_t i = 0; i < n; ++i) {
        double wn = (v[i] - oldeps * w2[i] - delta * w[i]) * ig;
        w2[i] = w[i];
        w[i] = wn;
        x[i] += phi * wn;
      }
      residual = phibar / beta1;
      if (residual <= rtol) {
        converged = 1;
        break;
      }
    }
    rep.iters = iters;
    rep.residual = residual;
    rep.converged = converged;
  }
done:
  free(r1);
  free(r2);
  free(y);
  free(v);
  free(w);
  free(w2);
  free(yn);
  return rep;
}

orc_report orc_minres(const orc_csr *a, const double *inv_diag, const double *b, double rtol,
                      int64_t max_iters, double *x) {
  precond pc;
  memset(&pc, 0, sizeof(pc));
  pc.inv_diag = inv_diag;
  return minres_core(a, &pc, b, rtol, max_iters, x);
}

orc_report orc_minres_cheb(const orc_csr *a, const orc_csr *p, const double *inv_diag,
                           int64_t split, int m, const double *coef, const double *b, double rtol,
                           int64_t max_iters, double *x) {
  precond pc;
  memset(&pc, 0, sizeof(pc));
  pc.inv_diag = inv_diag;
  pc.cheb_p = p;
  pc.coef = coef;
  pc.split = split;
  pc.m = m < 1 ? 1 : m;
  int64_t np = p->nrows;
  pc.res = (double *)malloc(sizeof(double) * (np ? np : 1));
  pc.d = (double *)malloc(sizeof(double) * (np ? np : 1));
  pc.dn = (double *)malloc(sizeof(double) * (np ? np : 1));
  pc.t = (double *)malloc(sizeof(double) * (np ? np : 1));
  orc_report rep = minres_core(a, &pc, b, rtol, max_iters, x);
  free(pc.res);
  free(pc.d);
  free(pc.dn);
  free(pc.t);
  return rep;
}
