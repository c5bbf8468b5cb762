/*
 * Exercises every entry point of the C oracle (oracle/feec_oracle.c) on small
 * Kuhn meshes in d = 1..4 so that an AddressSanitizer / UndefinedBehavior-
 * Sanitizer build reports any out-of-bounds access, leak, overflow or
 * misaligned load in the host code. Exit status 0 and the line
 * "oracle_check ok" mean every call returned and every result was finite.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "feec_oracle.h"

static int failures = 0;

static void expect(int ok, const char *what, int d, int k) {
  if (!ok) {
    fprintf(stderr, "FAIL %s (d=%d k=%d)\n", what, d, k);
    ++failures;
  }
}

static int finite_all(const double *x, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (!isfinite(x[i])) return 0;
  return 1;
}

static void check_csr(const orc_csr *a, const char *what, int d, int k) {
  int ok = a->row_ptr[0] == 0 && a->row_ptr[a->nrows] == a->nnz;
  for (int64_t i = 0; i < a->nrows; ++i) ok &= a->row_ptr[i] <= a->row_ptr[i + 1];
  for (int64_t p = 0; p < a->nnz; ++p) ok &= a->col[p] >= 0 && a->col[p] < a->ncols;
  ok &= finite_all(a->val, a->nnz);
  expect(ok, what, d, k);
}

static void mesh_case(int d, int grid) {
  int64_t nc = orc_kuhn_ncells(d, grid), nv = orc_kuhn_nverts(d, grid);
  int64_t *cells = malloc(sizeof(int64_t) * nc * (d + 1));
  double *coords = malloc(sizeof(double) * nv * d);
  orc_kuhn(d, grid, cells, coords);
  /* interior jitter so no entry is exactly representable */
  for (int64_t v = 0; v < nv * d; ++v) {
    double c = coords[v];
    if (c > 0.0 && c < 1.0) coords[v] = c + 0.01 * sin(17.0 * (double)v) / grid;
  }
  orc_complex *cx = orc_complex_new(d, nc, cells);
  expect(orc_dim(cx) == d, "dim", d, -1);
  int64_t ns[5];
  for (int k = 0; k <= d; ++k) {
    ns[k] = orc_nsimplices(cx, k);
    int64_t *sk = malloc(sizeof(int64_t) * ns[k] * (k + 1));
    orc_skeleton(cx, k, sk);
    int64_t nf = orc_binom(d + 1, k + 1);
    int64_t *cf = malloc(sizeof(int64_t) * nc * nf);
    orc_cell_faces(cx, k, cf);
    int ok = 1;
    for (int64_t i = 0; i < nc * nf; ++i) ok &= cf[i] >= 0 && cf[i] < ns[k];
    expect(ok, "cell_faces range", d, k);
    uint8_t *bf = malloc(ns[k] ? ns[k] : 1);
    orc_boundary_flags(cx, k, bf);
    int *loc = malloc(sizeof(int) * nf * (k + 1));
    expect(orc_local_subsets(d, k, loc) == nf, "local_subsets", d, k);
    free(loc);
    free(bf);
    free(cf);
    free(sk);
  }
  int8_t *signs = malloc(nc);
  expect(orc_orientation(cx, signs) == 1, "orientation", d, -1);
  free(signs);
  double *len = malloc(sizeof(double) * ns[1]);
  orc_lengths_from_coords(cx, coords, 0, len);
  expect(finite_all(len, ns[1]), "lengths", d, 1);
  double *lenm = malloc(sizeof(double) * ns[1]);
  orc_lengths_from_coords(cx, coords, 1, lenm);
  free(lenm);
  /* per-cell metric and element matrices of every pairing and grade */
  double g[16], gi[16], det;
  int nl = (d + 1) * d / 2;
  double *s = malloc(sizeof(double) * nl);
  for (int e = 0; e < nl; ++e) s[e] = 1.0 + 0.1 * e;
  orc_cell_metric(d, s, g, gi, &det);
  expect(isfinite(det), "cell_metric", d, -1);
  for (int kind = ORC_MASS; kind <= ORC_DIF_BOTH; ++kind)
    for (int k = 0; k <= d; ++k) {
      int r, c;
      orc_form_shape(d, kind, k, &r, &c);
      if (r <= 0 || c <= 0) continue;
      double *one = malloc(sizeof(double) * r * c);
      orc_element_matrix(d, kind, k, s, one);
      expect(finite_all(one, (int64_t)r * c), "element_matrix", d, k);
      free(one);
      double *all = malloc(sizeof(double) * nc * r * c);
      orc_element_matrices(cx, len, kind, k, all, 2);
      expect(finite_all(all, nc * r * c), "element_matrices", d, k);
      free(all);
      orc_csr *a = orc_assemble(cx, len, kind, k, 2);
      check_csr(a, "assemble", d, k);
      orc_csr_free(a);
    }
  free(s);
  /* mixed KKT of grade 1 (or 0 in 1D) from coboundary, masses and block algebra */
  int k = d >= 2 ? 1 : 0;
  orc_csr *mk = orc_assemble(cx, len, ORC_MASS, k, 1);
  orc_csr *kk = orc_assemble(cx, len, ORC_DIF_BOTH, k + 1, 1);
  orc_csr *cb = orc_coboundary(cx, k);
  check_csr(cb, "coboundary", d, k);
  orc_csr *cbt = orc_csr_transpose(cb);
  orc_csr *sum = orc_csr_add(mk, kk);
  check_csr(sum, "csr_add", d, k);
  orc_csr *prod = orc_csr_mul(cbt, cb);
  check_csr(prod, "csr_mul", d, k);
  uint8_t *keep = malloc(ns[k] ? ns[k] : 1);
  orc_boundary_flags(cx, k, keep);
  for (int64_t i = 0; i < ns[k]; ++i) keep[i] = !keep[i];
  orc_csr *sel = orc_csr_select(sum, keep, keep);
  check_csr(sel, "csr_select", d, k);
  double *sc = malloc(sizeof(double) * ns[k]);
  for (int64_t i = 0; i < ns[k]; ++i) sc[i] = 1.0 + (double)(i % 3);
  orc_csr *srows = orc_csr_scale_rows(sum, sc);
  check_csr(srows, "csr_scale_rows", d, k);
  const orc_csr *blocks[4] = {mk, NULL, NULL, kk};
  double scales[4] = {-1.0, 0.0, 0.0, 1.0};
  int64_t sizes[2] = {ns[k], ns[k]};
  orc_csr *blk = orc_csr_block(2, blocks, scales, sizes, sizes);
  check_csr(blk, "csr_block", d, k);
  /* solvers: CG on M+K (SPD), MINRES on the indefinite block, Chebyshev MINRES */
  int64_t n = sum->nrows;
  double *b = malloc(sizeof(double) * 2 * n), *x = calloc(2 * n, sizeof(double));
  double *idg = malloc(sizeof(double) * 2 * n);
  for (int64_t i = 0; i < 2 * n; ++i) b[i] = cos(0.37 * (double)i);
  for (int64_t i = 0; i < n; ++i) {
    double dg = 0.0;
    for (int64_t p = sum->row_ptr[i]; p < sum->row_ptr[i + 1]; ++p)
      if (sum->col[p] == i) dg = sum->val[p];
    idg[i] = idg[n + i] = 1.0 / dg;
  }
  orc_report rc = orc_cg(sum, idg, b, 1e-10, 10 * n + 10, x);
  expect(rc.converged, "cg", d, k);
  double *y = malloc(sizeof(double) * 2 * n);
  orc_spmv(sum, x, y);
  expect(finite_all(y, n), "spmv", d, k);
  memset(x, 0, sizeof(double) * 2 * n);
  orc_report rm = orc_minres(blk, NULL, b, 1e-8, 20 * n + 20, x);
  expect(isfinite(rm.residual), "minres", d, k);
  memset(x, 0, sizeof(double) * 2 * n);
  double coef[2] = {1.0, 1.0};
  orc_report rh = orc_minres_cheb(sum, sum, idg, n, 1, coef, b, 1e-8, 20 * n + 20, x);
  expect(isfinite(rh.residual), "minres_cheb", d, k);
  free(y);
  free(idg);
  free(x);
  free(b);
  free(sc);
  free(keep);
  orc_csr_free(blk);
  orc_csr_free(srows);
  orc_csr_free(sel);
  orc_csr_free(prod);
  orc_csr_free(sum);
  orc_csr_free(cbt);
  orc_csr_free(cb);
  orc_csr_free(kk);
  orc_csr_free(mk);
  free(len);
  orc_complex_free(cx);
  free(coords);
  free(cells);
}

int main(void) {
  const int sizes[5] = {0, 6, 4, 3, 2};
  for (int d = 1; d <= 4; ++d) mesh_case(d, sizes[d]);
  if (failures) {
    fprintf(stderr, "oracle_check: %d failures\n", failures);
    return 1;
  }
  printf("oracle_check ok\n");
  return 0;
}
