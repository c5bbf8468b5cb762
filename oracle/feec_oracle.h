/*
 * feec_oracle.h -- CPU FP64 restatement of formoniq's FEEC assembly-and-solve
 * path.
 *
 * TEST INFRASTRUCTURE. This library is the checker: only the test suite, the
 * smoke check and the CPU-baseline leg of the benchmark load it. The product
 * (libfeec, the HIP path) never links, loads or calls anything here.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the formoniq checkout, crates/...). All arithmetic is IEEE FP64.
 */
#ifndef FEEC_ORACLE_H
#define FEEC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Form kinds, as WhitneyPairing's four constructors (formoniq/src/operators.rs:163-193). */
enum { ORC_MASS = 0, ORC_DIF_TRIAL = 1, ORC_DIF_TEST = 2, ORC_DIF_BOTH = 3 };

typedef struct orc_csr {
  int64_t nrows, ncols, nnz;
  int64_t *row_ptr; /* nrows+1 */
  int64_t *col;     /* nnz */
  double *val;      /* nnz */
} orc_csr;

typedef struct orc_complex orc_complex;

/* ---- combinatorics / reference cell ---- */
int64_t orc_binom(int64_t n, int64_t k);
/* Local colex (k+1)-subsets of {0..d}: writes count*(k+1) vertex positions. */
int orc_local_subsets(int d, int k, int *out);
/* Element matrix shape of a pairing: rows, cols. */
void orc_form_shape(int d, int kind, int grade, int *rows, int *cols);

/* ---- meshers ---- */
/* Kuhn cells of an n^d box grid (generation order) and vertex coords [V][d]. */
int64_t orc_kuhn_ncells(int d, int n);
int64_t orc_kuhn_nverts(int d, int n);
void orc_kuhn(int d, int n, int64_t *cells, double *coords);

/* ---- topology ---- */
orc_complex *orc_complex_new(int d, int64_t ncells, const int64_t *cells);
void orc_complex_free(orc_complex *c);
int orc_dim(const orc_complex *c);
int64_t orc_nsimplices(const orc_complex *c, int k);
void orc_skeleton(const orc_complex *c, int k, int64_t *out);
void orc_cell_faces(const orc_complex *c, int k, int64_t *out);
void orc_boundary_flags(const orc_complex *c, int k, uint8_t *out);
orc_csr *orc_coboundary(const orc_complex *c, int k);
/* Coherent cell orientation signs (+1/-1), cell 0 positive; 0 if non-orientable. */
int orc_orientation(const orc_complex *c, int8_t *signs);

/* ---- geometry ---- */
void orc_lengths_from_coords(const orc_complex *c, const double *coords, int minkowski,
                             double *out);
/* Per-cell metric g [d*d], inverse [d*d], det. */
void orc_cell_metric(int d, const double *s, double *g, double *ginv, double *det);
/* Per-cell element matrix from the cell's local squared edge lengths. */
void orc_element_matrix(int d, int kind, int grade, const double *s, double *out);
void orc_element_matrices(const orc_complex *c, const double *lengths, int kind, int grade,
                          double *out, int nthreads);

/* ---- assembly ---- */
orc_csr *orc_assemble(const orc_complex *c, const double *lengths, int kind, int grade,
                      int nthreads);
orc_csr *orc_csr_new(int64_t nrows, int64_t ncols, int64_t nnz);
void orc_csr_free(orc_csr *a);
orc_csr *orc_csr_add(const orc_csr *a, const orc_csr *b);
orc_csr *orc_csr_mul(const orc_csr *a, const orc_csr *b);
orc_csr *orc_csr_transpose(const orc_csr *a);
orc_csr *orc_csr_select(const orc_csr *a, const uint8_t *row_keep, const uint8_t *col_keep);
orc_csr *orc_csr_scale_rows(const orc_csr *a, const double *s);
/* Block matrix from nb*nb blocks (NULL = empty), with scalars. */
orc_csr *orc_csr_block(int nb, const orc_csr *const *blocks, const double *scales,
                       const int64_t *row_sizes, const int64_t *col_sizes);
void orc_spmv(const orc_csr *a, const double *x, double *y);

/* ---- Krylov (iterative/src/krylov.rs) ---- */
typedef struct orc_report {
  int64_t iters;
  double residual;
  int converged;
} orc_report;
/* Preconditioner: NULL = identity, else y = inv_diag .* r (Jacobi / block Jacobi). */
orc_report orc_cg(const orc_csr *a, const double *inv_diag, const double *b, double rtol,
                  int64_t max_iters, double *x);
orc_report orc_minres(const orc_csr *a, const double *inv_diag, const double *b, double rtol,
                      int64_t max_iters, double *x);
/* MINRES with the Chebyshev block preconditioner: p is the block-diagonal
 * np x np hdif_gram matrix (rows < split: block 0), inv_diag its inverse
 * diagonal, coef = {1/theta_0, 1/theta_1, then (a_k0, b_k0, a_k1, b_k1) for
 * k < m-1}; rows >= np are left unpreconditioned. */
orc_report orc_minres_cheb(const orc_csr *a, const orc_csr *p, const double *inv_diag,
                           int64_t split, int m, const double *coef, const double *b, double rtol,
                           int64_t max_iters, double *x);

#ifdef __cplusplus
}
#endif
#endif
