/*
 * feec.h -- C ABI of libfeec, the MI355X (gfx950) implementation of formoniq's
 * FEEC assembly-and-solve hot path.
 *
 * Plain C: opaque handles, plain pointers and sizes. Every call returns a
 * feec_status; feec_last_error() gives a thread-local message for the last
 * failure on the calling thread. Host buffers passed in or out are
 * caller-owned and sized by the count queries. Handles are library-owned and
 * freed by their *_destroy function. Calls on one feec_ctx are serialized on
 * that context's HIP stream; use one context per host thread.
 *
 * Reference interfaces replaced (luiswirth/formoniq, paths relative to crates/):
 *   CartesianTopology::cell_skeleton      simplicial/src/mesher/grid.rs:77-98
 *   Complex::from_cells                    simplicial/src/topology/complex.rs:278-336
 *   Complex::boundary_simplices            simplicial/src/topology/complex.rs:140-149
 *   Complex::coboundary_operator           simplicial/src/topology/complex.rs:168-217
 *   MeshCoords::to_edge_lengths_sq         regge/src/coord/mesh.rs:208-216
 *   MeshLengthsSq::cell_metric             regge/src/lengths/mesh.rs:182-184
 *   BilinearForm::{element, assemble}      formoniq/src/galerkin.rs:42-58, 138-188
 *   WhitneyPairing::{mass,dif_trial,dif_test,dif_both}  formoniq/src/operators.rs:163-211
 *   HilbertComplex::{mass,...,hdif_gram,dif,inclusion}  formoniq/src/whitney_complex.rs:55-277
 *   elliptic::solve_source                 formoniq/src/problems/elliptic.rs:132-182
 *   HodgeDirac::assemble_signed            formoniq/src/problems/dirac.rs:213-258
 *   solve_dirac_source                     formoniq/src/problems/dirac.rs:473-534
 *   krylov::{cg, minres}                   iterative/src/krylov.rs:48-211
 */
#ifndef FEEC_H
#define FEEC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum feec_status {
  FEEC_OK = 0,
  FEEC_E_ARG = 1,             /* bad argument or shape (the reference panics) */
  FEEC_E_DEVICE = 2,          /* HIP runtime error */
  FEEC_E_OOM = 3,             /* device allocation failed */
  FEEC_E_TOPOLOGY = 4,        /* dangling/duplicate vertices, non-pseudomanifold */
  FEEC_E_DEGENERATE_CELL = 5, /* a cell metric with det g == 0 */
  FEEC_E_NOT_SPD = 6,         /* an SPD-only path met an indefinite block */
  FEEC_E_NOT_CONVERGED = 7,   /* Krylov budget spent; the report holds the state */
  FEEC_E_UNSUPPORTED = 8,     /* a case outside what the library implements */
  FEEC_E_ADMISSION = 9,       /* serving: request larger than the HBM budget */
  FEEC_E_AGAIN = 10           /* serving: result not ready yet */
} feec_status;

const char *feec_last_error(void);
const char *feec_status_string(feec_status s);

typedef struct feec_ctx_s feec_ctx;
typedef struct feec_mesh_s feec_mesh;
typedef struct feec_geom_s feec_geom;
typedef struct feec_csr_s feec_csr;
typedef struct feec_problem_s feec_problem;
typedef struct feec_server_s feec_server;

/* The four WhitneyPairing blocks (operators.rs:163-193); the grade argument is
 * always the grade of the inner product. FEEC_HDIF_GRAM is
 * HilbertComplex::hdif_gram(k) = mass(k) + dif_both(k+1) (whitney_complex.rs:180-183). */
typedef enum feec_form_kind {
  FEEC_MASS = 0,
  FEEC_DIF_TRIAL = 1,
  FEEC_DIF_TEST = 2,
  FEEC_DIF_BOTH = 3,
  FEEC_HDIF_GRAM = 4
} feec_form_kind;

/* Full complex (natural BC, WhitneyComplex) or relative complex (essential BC,
 * RelativeWhitneyComplex::new, whitney_complex.rs:545-557). */
typedef enum feec_bc { FEEC_FULL = 0, FEEC_RELATIVE = 1 } feec_bc;

typedef enum feec_ambient { FEEC_EUCLIDEAN = 0, FEEC_MINKOWSKI = 1 } feec_ambient;

typedef enum feec_precision {
  FEEC_PREC_FP64 = 0,      /* FP64 matrix, vectors and Krylov */
  FEEC_PREC_MIXED = 1,     /* FP32 SpMV/Krylov inside FP64 iterative refinement */
  FEEC_PREC_FP64_PC32 = 2  /* FP64 SpMV and Krylov; the Chebyshev polynomial runs in FP32 on the
                              FP16 SELL copy (the mixed solve's preconditioner, without its
                              restarts). Requires FEEC_PC_STORE_SELL */
} feec_precision;

/* Storage of the preconditioner matrix the Chebyshev polynomial runs on. */
typedef enum {
  FEEC_PC_STORE_SELL = 0,     /* D^-1/2 hdif_gram D^-1/2 as SELL-64 slices, off-diagonal entries
                                 with int16 column offsets (int32 columns when an offset exceeds
                                 +-32767), unit diagonal implied: FP16 values on the mixed path,
                                 FP64 on the FP64 path */
  FEEC_PC_STORE_CSR_HALF = 1, /* mixed path: D^-1/2 hdif_gram D^-1/2 in CSR with FP16 values;
                                 FP64 path: as FEEC_PC_STORE_CSR */
  FEEC_PC_STORE_CSR = 2       /* D^-1 hdif_gram in CSR with FP32 (mixed) or FP64 values */
} feec_pc_storage;

typedef struct feec_solve_opts {
  double rtol;           /* target true relative residual |b-Ax|/|b| (default 1e-10) */
  int64_t max_iters;     /* Krylov iteration budget over all refinements (default 200000) */
  int32_t precision;     /* feec_precision (default FEEC_PREC_FP64_PC32) */
  int32_t max_refinements; /* mixed: outer FP64 refinement steps (default 12) */
  double inner_rtol;     /* mixed: FP32 inner MINRES tolerance (default 1e-4) */
  int32_t use_graph;     /* capture Krylov iterations in a hipGraph (default 1) */
  int32_t graph_iters;   /* most iterations per graph segment, even (default 32); a segment is
                            cut to about 16 ms of estimated device work, so that other
                            contexts' work does not queue long behind it */
  int32_t cheb_degree;   /* degree of the Chebyshev-Jacobi polynomial applied to each
                            hdif_gram block of the preconditioner; 0 = Jacobi only; -1 = by
                            mesh resolution r = N_0^(1/d): 8 sqrt(r) in multiples of 8 within
                            [16, 128] (mixed) or half that within [24, 64] (FP64) (default -1) */
  double cheb_lower;     /* lower end of the Chebyshev interval as a fraction of its upper end
                            (1.05 lambda_max); 0 = by degree: 1e-2 up to degree 4, 3e-3 up to 8,
                            1.5e-3 up to 12, 1e-3 up to 32, 5e-4 up to 48, 2.5e-4 up to 64,
                            1.2e-4 up to 96, else 6e-5 (default 0) */
  int32_t precond_storage; /* feec_pc_storage: how the matrix the polynomial runs on is
                              stored (default FEEC_PC_STORE_SELL) */
  int32_t preconditioner; /* feec_preconditioner: the block preconditioner (default FEEC_PC_AUTO) */
  int32_t mg_sweeps;      /* Jacobi smoothing sweeps before and after the coarse correction on
                             each multigrid level (default 2) */
} feec_solve_opts;

/* The block preconditioner diag(B_{k-1}, B_k, I) of the mixed KKT system, B_j approximating
 * hdif_gram(j)^-1 (elliptic.rs:29-47).
 * FEEC_PC_MULTIGRID: the Hiptmair-Xu auxiliary-space preconditioner (hx.rs:150-275) with
 *   geometric V-cycles of the nodal problem (multigrid.rs:55-213): iteration counts bounded
 *   under refinement. It needs a Kuhn-grid mesh numbered as feec_kuhn_grid numbers it, a
 *   geometry from Euclidean coordinates and grade <= 2; otherwise FEEC_E_UNSUPPORTED.
 * FEEC_PC_CHEBYSHEV: a Chebyshev-Jacobi polynomial of degree cheb_degree on each block.
 * FEEC_PC_AUTO: multigrid where it applies and the system is not small (at least 20,000
 *   KKT rows), else Chebyshev. Batched solves (feec_batch_run) always run Chebyshev. */
typedef enum {
  FEEC_PC_AUTO = 0,
  FEEC_PC_CHEBYSHEV = 1,
  FEEC_PC_MULTIGRID = 2
} feec_preconditioner;

typedef struct feec_report {
  int64_t iters;           /* Krylov iterations over all refinements */
  int32_t refinements;     /* FP64 refinement steps taken (0 for plain FP64) */
  int32_t converged;       /* 1 when residual <= rtol */
  double residual;         /* true FP64 |b-Ax|_2/|b|_2 of the returned solution */
  double precond_residual; /* MINRES phibar/beta1 of the last Krylov run (the reference's stopping quantity) */
  double seconds;          /* wall time of the solve on the device */
} feec_report;

void feec_solve_opts_default(feec_solve_opts *opts);

/* ---- context --------------------------------------------------------- */
feec_status feec_ctx_create(int device, feec_ctx **out);
void feec_ctx_destroy(feec_ctx *ctx);
feec_status feec_ctx_synchronize(feec_ctx *ctx);
/* Device bytes currently held by handles of this context. */
feec_status feec_ctx_memory(const feec_ctx *ctx, int64_t *bytes_in_use);
/* Device allocations this context has made so far (blocks reused from its
 * cache included): a caller can check that a call allocated nothing. */
feec_status feec_ctx_allocations(const feec_ctx *ctx, int64_t *count);

/* ---- meshers (host) ---------------------------------------------------- */
/* CartesianGrid::new_unit(dim, n).triangulate() (regge/src/mesher/cartesian.rs:109-233):
 * Kuhn cells [ncells][dim+1] in generation order and coordinates [nverts][dim], axis 0 fastest. */
feec_status feec_kuhn_counts(int dim, int n, int64_t *ncells, int64_t *nverts);
feec_status feec_kuhn_grid(int dim, int n, int64_t *cells, double *coords);
/* Interior-vertex jitter (no reference counterpart; build-defined synthetic
 * input): each interior vertex of the n^dim unit grid, in vertex order, each axis
 * a with bit a of axis_mask set, gets amplitude*h*(2U-1), U from SplitMix64(seed). */
feec_status feec_jitter_interior(int dim, int n, double *coords, double amplitude, uint64_t seed,
                                 uint32_t axis_mask);
/* N(0,1) values from SplitMix64(seed) by Box-Muller (build-defined). */
feec_status feec_normal_vector(int64_t n, uint64_t seed, double *out);

/* ---- topology ---------------------------------------------------------- */
/* Complex::from_cells: colex numbering of every skeleton, built on the device
 * by sort. Cells: [ncells][dim+1], vertices 0..nverts-1, all used. */
feec_status feec_mesh_create(feec_ctx *ctx, int dim, int64_t nverts, int64_t ncells,
                             const int64_t *cells, feec_mesh **out);
void feec_mesh_destroy(feec_mesh *mesh);
feec_status feec_mesh_dim(const feec_mesh *mesh, int *dim);
feec_status feec_mesh_counts(const feec_mesh *mesh, int64_t *counts /* dim+1 */);
feec_status feec_mesh_skeleton(const feec_mesh *mesh, int k, int64_t *verts /* N_k*(k+1) */);
feec_status feec_mesh_cell_faces(const feec_mesh *mesh, int k, int64_t *out /* N_d*C(d+1,k+1) */);
feec_status feec_mesh_boundary(const feec_mesh *mesh, int k, uint8_t *flags /* N_k */);
/* Coherent orientation signs of the cells (cell 0 positive); FEEC_E_TOPOLOGY if non-orientable. */
feec_status feec_mesh_orientation(feec_mesh *mesh, int8_t *signs /* N_d */);
/* HilbertComplex::dif(k): the coboundary D_k as a CSR (rows N_{k+1}, cols N_k). */
feec_status feec_mesh_coboundary(feec_mesh *mesh, feec_bc bc, int k, feec_csr **out);

/* ---- geometry ---------------------------------------------------------- */
feec_status feec_geom_from_lengths(feec_mesh *mesh, const double *lengths_sq /* N_1 */,
                                   feec_geom **out);
feec_status feec_geom_from_coords(feec_mesh *mesh, const double *coords /* nverts*dim */,
                                  feec_ambient ambient, feec_geom **out);
void feec_geom_destroy(feec_geom *geom);
feec_status feec_geom_lengths(const feec_geom *geom, double *out /* N_1 */);
/* Per cell: g [d*d], g^-1 [d*d], det g; any pointer may be NULL. */
feec_status feec_cell_metrics(feec_geom *geom, double *g, double *ginv, double *det);
/* Signature guards (metric/src/lib.rs:47-58, 203-205; regge/src/lengths/mesh.rs:219-236):
 * edges with s<0, s==0, s>0, and per-cell Riemannian flags (Cholesky succeeds). */
feec_status feec_causal_census(feec_geom *geom, int64_t counts[3], uint8_t *cell_riemannian);

/* ---- element matrices and assembly ------------------------------------- */
/* Row-major [N_d][rows][cols] in local colex face order. */
feec_status feec_element_shape(int dim, feec_form_kind kind, int grade, int *rows, int *cols);
feec_status feec_element_matrices(feec_geom *geom, feec_form_kind kind, int grade, double *out);
/* The Galerkin matrix: element entries evaluated with the reference's
 * operations in its order, so pattern, numbering, signs and values are
 * bit-exact with the CPU restatement of the reference (exact-zero element
 * entries dropped, duplicates summed in cell order, summed zeros kept);
 * relative = E^T A E. */
feec_status feec_assemble(feec_geom *geom, feec_bc bc, feec_form_kind kind, int grade,
                          feec_csr **out);
feec_status feec_csr_shape(const feec_csr *a, int64_t *nrows, int64_t *ncols, int64_t *nnz);
feec_status feec_csr_copy(const feec_csr *a, int64_t *row_ptr, int32_t *col, double *val);
feec_status feec_csr_spmv(feec_csr *a, const double *x, double *y); /* host vectors */
/* Views of a library CSR's arrays (int64 row_ptr, int32 col, FP64 val), valid until
 * feec_csr_destroy. on_device = 1: the device arrays themselves, in the context's memory
 * (the context's stream is drained first, so device code of the caller may read them);
 * on_device = 0: a library-owned host copy made on the first request. Any pointer may be NULL. */
feec_status feec_csr_view(const feec_csr *a, int64_t *nrows, int64_t *ncols, int64_t *nnz,
                          const int64_t **row_ptr, const int32_t **col, const double **val,
                          int on_device);
/* The device-vector seam of iterative/src/lib.rs:58-186 (LinearOperator::apply,
 * InnerProductSpace::{dot, add_scaled}) on caller-owned device arrays of the matrix's
 * context's device; each call is complete on return. */
feec_status feec_csr_spmv_device(const feec_csr *a, const double *x, double *y); /* y = A x */
feec_status feec_vec_dot_device(feec_ctx *ctx, int64_t n, const double *x, const double *y,
                                double *out); /* *out (host) = x . y, fixed summation order */
feec_status feec_vec_axpy_device(feec_ctx *ctx, int64_t n, double a, const double *x,
                                 double *y); /* y += a x */
/* The Galerkin load of a cochain, b = M_k c (N_k entries, full complex): the L^2
 * functional of its Whitney interpolant, galerkin.rs:324-371, the type-correct way to pass
 * a cochain to feec_solve_source. M_k is assembled as feec_assemble does; the product sums
 * each row in column order (nalgebra-sparse's CSR x dense). */
feec_status feec_load_vector(feec_geom *geom, int grade, const double *cochain, double *out);
/* krylov::cg with the Jacobi preconditioner (krylov.rs:48-95, precond.rs:53-123) on a library
 * CSR, scalars on the device: x (host) from x0 = 0 until |r| <= rtol |b|. The report holds the
 * iterations and |r|/|b|; FEEC_E_NOT_CONVERGED when max_iters ends the run first. */
feec_status feec_csr_cg(feec_csr *a, const double *b, double *x, double rtol, int64_t max_iters,
                        feec_report *report);
/* feec_csr_cg without any device allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same kernels,
 * same order, same results as feec_csr_cg; its vectors, partial sums and scalars
 * live in a file-scope device array set up by feec_ctx_create and held by each GPU
 * until the process ends (14.1 MiB: 262144 rows at most; larger a: FEEC_E_ARG).
 * Every call on one GPU shares that array, so calls that overlap on one GPU
 * (from several threads or contexts) race; serialize them. */
feec_status feec_csr_cg_noalloc(feec_csr *a, const double *b, double *x, double rtol,
                                int64_t max_iters, feec_report *report);
void feec_csr_destroy(feec_csr *a);

/* ---- analytic data: quadrature loads, L2 projection and error, de Rham map ------------
 * A differential form reaches the library as its values at quadrature nodes, evaluated by
 * the caller (on the host or on the device) at the positions that
 * feec_cell_quadrature_points / feec_simplex_quadrature_points report. A k-form value is its
 * C(dim,k) components on the colex basis dx^I. frame FEEC_FRAME_REFERENCE: components in each
 * cell's reference frame, a Section's value (derham/src/section.rs:67); FEEC_FRAME_AMBIENT:
 * components in the coordinates the geometry was made from (feec_geom_from_coords), pulled
 * back cell by cell as Pullback::at does (section.rs:182-195). on_device = 1: every array
 * argument is device memory of the geometry's context, else host memory. Full complex only.
 * Arrays: cell nodes [N_dim][npoints][...], simplex nodes [N_k][npoints_k][...]. */
typedef enum feec_frame { FEEC_FRAME_REFERENCE = 0, FEEC_FRAME_AMBIENT = 1 } feec_frame;

/* SimplexQuadRule::degree (simplicial/src/atlas/quadrature.rs:37-83): the Grundmann-Moller
 * rule of index degree/2 on the dim-simplex (0 <= dim <= 4, 0 <= degree <= 7), barycentric
 * nodes [npoints][dim+1] and weights summing to one. bary and weights may be NULL. */
feec_status feec_quadrature_rule(int dim, int degree, int *npoints, double *bary,
                                 double *weights);
/* The embedded position x = sum_i lambda_i x_(v_i) of every cell's nodes: [N_dim][np][dim]. */
feec_status feec_cell_quadrature_points(feec_geom *geom, int degree, double *points,
                                        int on_device);
/* The same for the nodes of the rule on the k-simplices: [N_k][np_k][dim]. */
feec_status feec_simplex_quadrature_points(feec_geom *geom, int grade, int degree,
                                           double *points, int on_device);
/* SourceForm (operators.rs:599-635) assembled into its load vector (galerkin.rs:205-312):
 * load_sigma = sum_K vol_K sum_n w_n <f(x_n), W_sigma(x_n)>_(Lambda^k g_K^-1), N_k entries,
 * summed over cells in cell order; field at the cell nodes of `degree`. */
feec_status feec_source_load(feec_geom *geom, int grade, int degree, feec_frame frame,
                             const double *field, double *load, int on_device);
/* l2_projection (fe.rs:87-123): M_k c = load, by Jacobi CG to rtol 1e-12 when every cell is
 * Riemannian, else by MINRES (the reference factorizes). mass: the grade-k mass from
 * feec_assemble(geom, FEEC_FULL, FEEC_MASS, grade), reused across calls, or NULL to assemble
 * it inside the call as the reference does. report may be NULL. */
feec_status feec_l2_projection(feec_geom *geom, int grade, int degree, feec_frame frame,
                               const double *field, const feec_csr *mass, double *cochain,
                               feec_report *report, int on_device);
/* WhitneyInterpolant::eval (derham/src/interpolate/interpolant.rs:55-62): the reference-frame
 * value of W c at every cell's nodes of `degree`, [N_dim][np][C(dim,k)]. */
feec_status feec_whitney_at_nodes(feec_geom *geom, int grade, int degree, const double *cochain,
                                  double *values, int on_device);
/* fe_l2_error (fe.rs:50-74): |exact - W c|_(L^2 Lambda^k) by the degree-3 rule, measured in
 * each cell's frame by Lambda^k g^-1; exact at the degree-3 cell nodes. *error is a host double. */
feec_status feec_l2_error(feec_geom *geom, int grade, const double *cochain, feec_frame frame,
                          const double *exact, double *error, int on_device);
/* derham_map (derham/src/project.rs:55-96): c_s = integral over the k-simplex s of the form,
 * the pairing with its tangent blade by the degree-`degree` rule on the simplex; field in
 * ambient components at the simplex nodes [N_k][np_k][C(dim,k)]. */
feec_status feec_derham_map(feec_geom *geom, int grade, int degree, const double *field,
                            double *cochain, int on_device);

/* ---- request ingestion: gmsh and CBOR --------------------------------------------------
 * Host functions on byte buffers. Each reader is called twice: with the output arrays NULL
 * it reports the sizes, then it fills them. A file that is not what it claims is FEEC_E_ARG
 * with the reason in feec_last_error(), never a crash. */
/* gmsh2coord_cells_ordered (regge/src/io/gmsh.rs:84-175): MSH 4.1 ASCII. dim: the highest
 * simplex dimension present (0..3; its elements are the cells). cells [ncells][dim+1]: each
 * element's nodes in the order the file lists them, renumbered to the nodes the cells use by
 * increasing tag (node tag t is node t-1 of the file's node list, as the reference reads it).
 * coords [nverts][coord_dim]: coord_dim 2 when z is 0 on every node, else 3. Binary MSH and
 * parametric node blocks are FEEC_E_UNSUPPORTED. */
feec_status feec_gmsh_read(const char *bytes, int64_t len, int *dim, int *coord_dim,
                           int64_t *nverts, int64_t *ncells, int64_t *cells, double *coords);
/* Complex::load (simplicial/src/topology/complex.rs:91-99): the CBOR of the cell skeleton,
 * an array of {"vertices": [...]} (serde through ciborium 0.2.2). cells [ncells][dim+1]. */
feec_status feec_cbor_read_cells(const uint8_t *bytes, int64_t len, int *dim, int64_t *ncells,
                                 int64_t *cells);
/* MeshLengthsSq::load (regge/src/lengths/mesh.rs:244-251): {"vector": {"data": [...],
 * "nrows": n, "ncols": null}} (a plain array of numbers as the vector is read too). */
feec_status feec_cbor_read_lengths(const uint8_t *bytes, int64_t len, int64_t *nedges,
                                   double *lengths_sq);
/* The writers of the same encodings (floats in the shortest exact IEEE width, as ciborium
 * writes them). *len: the buffer's size in, the encoding's size out; out NULL to query. */
feec_status feec_cbor_write_cells(int dim, int64_t ncells, const int64_t *cells, uint8_t *out,
                                  int64_t *len);
feec_status feec_cbor_write_lengths(int64_t nedges, const double *lengths_sq, uint8_t *out,
                                    int64_t *len);

/* ---- autotuning (no counterpart in the reference) --------------------------------------
 * Launch choices that change no result bit are tuned per device and per (kernel family, dim,
 * grade, size class ceil(log2 n)) by timing every candidate on the first call that needs
 * them; the table lives as long as the process. Tuned now: the element kernels of the KKT
 * and Hodge-Dirac blocks fused in one launch or split into one launch per grade, and the
 * persistent grid of the batched element launch. */
/* mode 0: tune (default); mode c >= 1: always take candidate c-1 (tests, comparisons). */
feec_status feec_tune_mode(int mode);
/* The table as text, one line per entry: the key, the choice and each candidate's median
 * time in ms. *len: the buffer's size in, the text's size with its NUL out; out NULL to query. */
feec_status feec_tune_table(char *out, int64_t *len);
/* Sustained HBM rates of this device with plain streaming kernels over `bytes`-sized buffers
 * (16 bytes per lane, non-temporal writes, best of 5): rates[0] write, rates[1] read,
 * rates[2] copy (bytes read + written), in GB/s. The achievable ceiling beside the published
 * 8 TB/s peak when reading a kernel's roofline fraction. */
feec_status feec_ctx_stream_rates(feec_ctx *ctx, int64_t bytes, double rates[3]);

/* ---- problems ------------------------------------------------------------ */
/* b_k and an M_k-orthonormal harmonic basis [count][ndofs] (basis may be NULL): the
 * reading solve_harmonics returns (elliptic.rs:197-220). Kuhn boxes take the closed forms
 * (full k=0: the constant; relative k=n: M_n^-1 times the orientation; b_k = 0 otherwise).
 * Other meshes: b_k from exact ranks, then the projected integral cocycles of
 * harmonics (harmonic.rs:118-175), or the same closed forms where they are the whole
 * space on a connected mesh. FEEC_E_UNSUPPORTED: more than 8 harmonic forms, an
 * indefinite L2 pairing on the harmonic space (the reference's eigensolve fallback is
 * not built), or exact elimination beyond its budget. */
feec_status feec_harmonics(feec_geom *geom, feec_bc bc, int grade, int64_t *count, double *basis);
/* The period-normalized harmonic basis [count][ndofs] (Harmonics::integral,
 * harmonic.rs:118-175): column i pairs to delta_ij with the j-th integral cycle of
 * feec_homology_generators. */
feec_status feec_harmonics_integral(feec_geom *geom, feec_bc bc, int grade, int64_t *count,
                                    double *basis);
/* Betti number b_k of the complex (FEEC_FULL; Complex::betti_number, homology.rs:47-52)
 * or relative to its boundary (FEEC_RELATIVE; relative_betti_number, homology.rs:81-98):
 * n_k - rank d_k - rank d_{k+1}, ranks over two prime fields (exact.rs:97-147). */
feec_status feec_betti_number(const feec_mesh *mesh, feec_bc bc, int grade, int64_t *b);
/* Integral representatives of a basis of the free part of H^k (cocycles;
 * cohomology.rs:55-90) or of H_k (cycles; homology.rs:116-153) on the complex's DOFs
 * (relative: interior numbering), [count][ndofs] (out may be NULL to query count): the
 * RREF kernel vectors kept by quotient_generators (exact.rs:178-264), primitive, first
 * nonzero entry positive. */
feec_status feec_cohomology_generators(const feec_mesh *mesh, feec_bc bc, int grade,
                                       int64_t *count, int64_t *out);
feec_status feec_homology_generators(const feec_mesh *mesh, feec_bc bc, int grade,
                                     int64_t *count, int64_t *out);
/* The symmetric mixed KKT (elliptic.rs:54-114), with its harmonic border and
 * the sigma-row sign flip; sizes = {n_sigma, n_u, n_harmonic}. */
feec_status feec_mixed_kkt(feec_geom *geom, feec_bc bc, int grade, feec_csr **out,
                           int64_t sizes[3]);
/* elliptic::solve_source: source_galvec has N_k entries (ambient); sigma (N_{k-1}),
 * u (N_k) are ambient cochains, p has n_harmonic entries. */
feec_status feec_solve_source(feec_geom *geom, feec_bc bc, int grade, const double *source_galvec,
                              const feec_solve_opts *opts, double *sigma, double *u, double *p,
                              feec_report *report);
/* HodgeDirac::assemble_signed(sign): mass_block, op and offsets[dim+2]. */
feec_status feec_hodge_dirac_assemble(feec_geom *geom, feec_bc bc, double sign,
                                      feec_csr **mass_block, feec_csr **op, int64_t *offsets);
/* solve_dirac_source on the relative complex: (A + m M) u = f with lift;
 * load and lift are flat full-complex fields (lift NULL: zero), out is the flat
 * full field lift + E u_rel. Non-finite load, lift or mass_term: FEEC_E_ARG. */
feec_status feec_solve_dirac_source(feec_geom *geom, double mass_term, const double *load,
                                    const double *lift, const feec_solve_opts *opts, double *out,
                                    feec_report *report);
/* Leapfrog::step / conserved_energy (time.rs:292-420) on HodgeDirac::assemble of the complex,
 * 2-colored by grade parity (dirac.rs:330-337), as solve_dirac_leapfrog's loop
 * (dirac.rs:415-434): y (the flat field of the complex, offsets as feec_hodge_dirac_assemble)
 * advances nsteps steps of dt in place; states ((nsteps + 1) * ndofs values, or null) receives
 * every state from the initial one on, as the reference's returned vector of fields; energy
 * (nsteps + 1 values, or null) the staggered invariant of each. Mass solves by Jacobi CG to
 * 1e-14 (the reference factors them), so the mass matrix must be SPD (a Riemannian geometry;
 * otherwise FEEC_E_NOT_CONVERGED); the report counts CG iterations and device seconds. */
feec_status feec_dirac_leapfrog(feec_geom *geom, feec_bc bc, double dt, int64_t nsteps, double *y,
                                double *states, double *energy, feec_report *report);
/* LinearIrk::step with zero forcing (time.rs:191-274) and Tableau::gauss_legendre(stages),
 * stages 1 or 2 (solve_dirac, dirac.rs:383-403, uses 2), on HodgeDirac::assemble of the
 * complex: y, states and energy as feec_dirac_leapfrog, energy being 1/2 y^T M y. The stage
 * matrix I (x) M - dt A_tab (x) A is solved per step by restarted GMRES to 1e-14 (the reference
 * factors it with LU); the report counts GMRES iterations and device seconds. */
feec_status feec_dirac_irk(feec_geom *geom, feec_bc bc, int stages, double dt, int64_t nsteps,
                           double *y, double *states, double *energy, feec_report *report);

/* ---- resident problems (device-resident state, pattern reused) ----------- */
feec_status feec_problem_create(feec_geom *geom, feec_bc bc, int grade, const feec_solve_opts *opts,
                                feec_problem **out);
void feec_problem_destroy(feec_problem *prob);
/* Recompute element matrices and CSR values from the geometry (pattern kept). */
feec_status feec_problem_assemble(feec_problem *prob);
/* Solve with the source already on the device (N_k doubles); solution stays
 * resident; feec_problem_solution copies sigma/u/p out. */
feec_status feec_problem_solve_device(feec_problem *prob, const double *source_device,
                                      feec_report *report);
feec_status feec_problem_solve(feec_problem *prob, const double *source_host, feec_report *report);
/* Solve with a cochain source c (N_k doubles, device or host): the Galerkin load b = M_k c
 * (as feec_load_vector, bit for bit) is formed on the device from the problem's own
 * element blocks, then solved as feec_problem_solve_device(b). The first call builds the
 * full-complex M_k plan (kept with the problem). */
feec_status feec_problem_solve_cochain_device(feec_problem *prob, const double *cochain_device,
                                              feec_report *report);
feec_status feec_problem_solve_cochain(feec_problem *prob, const double *cochain_host,
                                       feec_report *report);
feec_status feec_problem_solution(feec_problem *prob, double *sigma, double *u, double *p);
feec_status feec_problem_sizes(const feec_problem *prob, int64_t sizes[3], int64_t *nnz);
/* Preconditioner matrix as a solve uses it (feec_problem_preconditioner). */
typedef enum {
  FEEC_PC_CSR_F32 = 0,    /* D^-1 hdif_gram, FP32 values, CSR */
  FEEC_PC_CSR_F16 = 1,    /* D^-1/2 hdif_gram D^-1/2, FP16 values, CSR */
  FEEC_PC_SELL64_F16 = 2, /* the same, off-diagonal only, packed {FP16, int16 offset} in 64-row
                             SELL slices */
  FEEC_PC_CSR_F64 = 3,    /* D^-1 hdif_gram, FP64 values, CSR */
  FEEC_PC_SELL64_F64 = 4, /* D^-1/2 hdif_gram D^-1/2, off-diagonal only, FP64 values and int16
                             offsets in 64-row SELL slices */
  FEEC_PC_SELL64_F16_C32 = 5, /* as FEEC_PC_SELL64_F16 with FP16 values and int32 columns apart
                                 (some column offset exceeds +-32767) */
  FEEC_PC_SELL64_F64_C32 = 6, /* as FEEC_PC_SELL64_F64 with int32 columns */
  FEEC_PC_MULTIGRID_F32 = 7,  /* the multigrid / HX blocks (FEEC_PC_MULTIGRID): V-cycles on FP32
                                 Kuhn-stencil diagonals, Chebyshev steps in FP32 */
  FEEC_PC_MULTIGRID_F64 = 8   /* the same in FP64 (the FP64 path) */
} feec_pc_format;

/* The block preconditioner blockdiag(hdif_gram(k-1), hdif_gram(k)) as built:
 * its nonzeros, the Chebyshev degree in use (0 = Jacobi; with the multigrid
 * preconditioner, the B applications per block application), the estimated
 * largest eigenvalue of D^-1 hdif_gram (with the multigrid preconditioner, of
 * B hdif_gram) per block, the bytes per stored CSR
 * column index (2 when every |col - row| fits a 16-bit offset, else 4), the
 * format a solve with the problem's options applies (feec_pc_format), the
 * entries that format stores (with SELL padding) and the off-diagonal
 * nonzeros; any pointer may be NULL. */
feec_status feec_problem_preconditioner(const feec_problem *prob, int64_t *nnz,
                                        int32_t *cheb_degree, double lambda_max[2],
                                        int32_t *column_bytes, int32_t *format, int64_t *stored,
                                        int64_t *offdiag_nnz);
/* Diagnostics (no counterpart in the reference; the reference applies its
 * preconditioner inside MINRES, krylov.rs): z = B r for host vectors r, z of
 * the KKT order (sigma, u, harmonic rows; sizes[0] + sizes[1] + sizes[2]
 * doubles), with B the block preconditioner exactly as a solve of `precision`
 * with `storage` (feec_pc_storage) applies it. */
feec_status feec_problem_apply_preconditioner(feec_problem *prob, int32_t precision,
                                              int32_t storage, const double *r, double *z);
/* Device pointer to an N_k scratch vector for feec_problem_solve_device. */
feec_status feec_problem_source_buffer(feec_problem *prob, double **device_ptr);
/* Copy an N_k host source into that device buffer. */
feec_status feec_problem_upload_source(feec_problem *prob, const double *source_host);
/* Diagnostics (no counterpart in the reference): average seconds per launch,
 * from HIP events on the context stream, of {element kernel, KKT value
 * scatter, fused MINRES SpMV in FP32, the same in FP64, Chebyshev step (the
 * preconditioner's fused SpMV, with the problem's precond_storage) in FP32,
 * the same in FP64}; the last four replayed from a graph of `reps` launches,
 * as a solve replays them. */
feec_status feec_problem_kernel_times(feec_problem *prob, int reps, double out[6]);
/* Diagnostics (no counterpart in the reference) for a problem whose
 * preconditioner is the multigrid one (else FEEC_E_UNSUPPORTED): average
 * seconds per launch, replayed from a graph of `reps` launches on the context
 * stream (use_graph != 0) or launched one by one, of {a finest-level Jacobi sweep (FP32), one block preconditioner
 * application as an FP64_PC32 solve applies it, one product by the
 * preconditioner matrix in the Chebyshev steps (FP32; each launch timed on
 * its own events right after an HX application, as a Chebyshev step meets it),
 * one HX application B (FP32)}. */
feec_status feec_problem_mg_kernel_times(feec_problem *prob, int reps, int use_graph, double out[4]);

/* Diagnostics (no counterpart in the reference): average seconds per launch of
 * the all-grade Hodge-Dirac element kernel -- M_k for k = 0..dim and
 * dif_trial(k) for k = 1..dim, every cell, one launch -- and the element
 * values it writes per cell. */
feec_status feec_dirac_element_time(feec_geom *geom, int reps, double *seconds,
                                    int64_t *values_per_cell);

/* ---- batching and serving -------------------------------------------------- */
/* One request: a jittered Kuhn box and a Hodge-Laplace source solve. */
typedef struct feec_request {
  int32_t dim;          /* 2..4 */
  int32_t grade;        /* 0..dim */
  int32_t n;            /* cells per axis */
  int32_t bc;           /* feec_bc */
  double jitter;        /* amplitude in units of h */
  uint64_t mesh_seed;   /* jitter seed */
  uint64_t source_seed; /* N(0,1) source cochain seed */
  int32_t precision;    /* feec_precision */
  int32_t reserved;
} feec_request;

typedef struct feec_result {
  int64_t ticket;
  int64_t ndofs;        /* KKT size */
  int64_t ncells;
  int64_t iters;
  int32_t refinements;
  int32_t status;       /* feec_status of this request */
  double residual;      /* true relative residual */
  double u_norm;        /* |u|_2 of the solution (for alone-vs-mixed checks) */
  double u_checksum;    /* sum_i u_i * ((i % 7) - 3) */
  double admit_time;    /* seconds since server start (batch: since the call, when every
                           mesh, plan and element block was done) */
  double finish_time;
  double latency;       /* finish - submit, seconds */
  double build_seconds; /* this request's mesh, numbering, plans and assembly (server) or
                           mesh, numbering and plans (batch) */
  double solve_seconds; /* its Krylov solve with refinement (batch: the lockstep solve of the
                           whole batch) */
} feec_result;

/* Batched: all element matrices of all requests in one persistent launch over
 * a (mesh, cell-range) work queue, then the solves with their MINRES runs in
 * lockstep (one launch per phase for every request still running); results
 * equal single solves bit for bit. */
feec_status feec_batch_run(feec_ctx *ctx, int nreq, const feec_request *reqs,
                           const feec_solve_opts *opts, feec_result *results);

typedef struct feec_server_opts {
  int64_t hbm_budget_bytes; /* admission budget for resident state */
  int32_t max_inflight;     /* slots: requests resident on the device at once */
  int32_t reserved;
  feec_solve_opts solve;
} feec_server_opts;

feec_status feec_server_create(feec_ctx *ctx, const feec_server_opts *opts, feec_server **out);
feec_status feec_server_submit(feec_server *srv, const feec_request *req, int64_t *ticket);
/* FEEC_E_AGAIN while the request is queued or running. */
feec_status feec_server_poll(feec_server *srv, int64_t ticket, feec_result *result);
/* Copy the solution u (N_k) of a finished request; NULL-safe size query. */
feec_status feec_server_solution(feec_server *srv, int64_t ticket, double *u, int64_t *n);
feec_status feec_server_drain(feec_server *srv);
/* Forget a finished request (its result and host copy of u): a long stream
 * holds only what its caller has not released. FEEC_E_AGAIN while it runs. */
feec_status feec_server_release(feec_server *srv, int64_t ticket);
/* Requests the server holds (queued, running or finished and not released)
 * and the doubles of solution held on the host for them. */
feec_status feec_server_held(feec_server *srv, int64_t *entries, int64_t *solution_doubles);
void feec_server_destroy(feec_server *srv);

#ifdef __cplusplus
}
#endif
#endif /* FEEC_H */
