// Internal declarations shared by the libfeec translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/feec.h"

namespace feec {

// ---------------------------------------------------------------------------
// errors: exceptions inside, feec_status at the C boundary
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  feec_status status;
  Error(feec_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

void set_last_error(const std::string& msg);

[[noreturn]] inline void fail(feec_status s, const std::string& m) { throw Error(s, m); }

#define FEEC_HIP(call)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      (void)hipGetLastError();                                                              \
      ::feec::fail(e_ == hipErrorOutOfMemory ? FEEC_E_OOM : FEEC_E_DEVICE,                  \
                   std::string(#call) + ": " + hipGetErrorString(e_));                      \
    }                                                                                       \
  } while (0)

// Enumerations cross the C ABI as ints, so a caller can pass any value, and in
// C++ reading an enum object that holds a value outside its declared range is
// undefined. Entry points therefore read such an argument's bits (no enum
// load) and check them against the declared values before anything else
// reads the argument.
template <class E>
inline int abi_int(const E& e) {
  static_assert(sizeof(E) == sizeof(int), "C enumerations are int-sized");
  int v;
  std::memcpy(&v, &e, sizeof v);
  return v;
}
inline bool abi_bc_ok(const feec_bc& b) {
  const int v = abi_int(b);
  return v == FEEC_FULL || v == FEEC_RELATIVE;
}

#define FEEC_REQUIRE(cond, status, msg) \
  do {                                      \
    if (!(cond)) ::feec::fail(status, msg); \
  } while (0)

// Every C-ABI entry runs in guard(). HIP keeps a per-thread "last error" that
// library code (rocPRIM) reads after its launches; a failure already reported,
// or one of a cleanup call whose status is ignored, is cleared on entry so it
// is never attributed to an unrelated later call.
template <class F>
feec_status guard(F&& f) {
  (void)hipGetLastError();
  try {
    f();
    return FEEC_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return FEEC_E_OOM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return FEEC_E_ARG;
  }
}

// ---------------------------------------------------------------------------
// small combinatorics shared by host and device
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  int r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

// ---------------------------------------------------------------------------
// device memory
// ---------------------------------------------------------------------------
struct Ctx;

void* device_alloc(Ctx* ctx, size_t bytes);
void device_free(Ctx* ctx, void* p, size_t bytes);

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  Ctx* ctx = nullptr;
  DBuf() = default;
  DBuf(Ctx* c, size_t count) { alloc(c, count); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n), ctx(o.ctx) {
    o.p = nullptr;
    o.n = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      reset();
      p = o.p;
      n = o.n;
      ctx = o.ctx;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  void alloc(Ctx* c, size_t count) {
    reset();
    ctx = c;
    n = count;
    p = static_cast<T*>(device_alloc(c, (count ? count : 1) * sizeof(T)));
  }
  void reset() {
    if (p) device_free(ctx, p, (n ? n : 1) * sizeof(T));
    p = nullptr;
    n = 0;
  }
  ~DBuf() { reset(); }
  size_t bytes() const { return n * sizeof(T); }
};

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::atomic<int64_t> bytes{0};
  std::atomic<int64_t> allocs{0};  // device_alloc calls (cache hits included)
  // rocPRIM temporary storage, grown on demand
  void* temp = nullptr;
  size_t temp_bytes = 0;
  void* ensure_temp(size_t bytes);
  // small device scalars owned by the context: an int flag and a counter
  int* dflag = nullptr;
  unsigned long long* dcount = nullptr;
  // pinned host words for reading solver scalars back, owned by the context
  void* pinned = nullptr;
  static constexpr size_t kPinnedBytes = 4096;
  // device blocks this context freed, by size class, for its own next
  // allocations (reuse in the stream's order); handed to the device-wide
  // cache when the context is destroyed
  std::map<size_t, std::vector<void*>> cache;
  void sync() { FEEC_HIP(hipStreamSynchronize(stream)); }
};
// Return a destroyed context's cached blocks (its stream drained) to the
// device-wide cache.
void release_cache(Ctx* ctx);

// Copies ordered on the context's stream; the host buffer may be reused on
// return (the stream is synchronized). Device memory is allocated and freed
// stream-ordered on the same stream, so no copy may go through another stream.
void copy_to_device(Ctx* c, void* dst, const void* src, size_t bytes);
void copy_to_host(Ctx* c, void* dst, const void* src, size_t bytes);

// Device COO -> CSR plan: the sorted order of (row, col) keys, kept so that a
// later value pass is a segmented sum (duplicates summed in input order).
// A COO source index with this bit set contributes its value negated.
constexpr uint32_t kCooNegate = 0x80000000u;

struct CooPlan {
  int64_t nrows = 0, ncols = 0;
  int64_t nnz = 0;        // unique (row, col) entries
  DBuf<int64_t> row_ptr;  // nrows+1
  DBuf<int32_t> col;      // nnz
  DBuf<int64_t> seg;      // nnz+1: range of each entry in `src`
  DBuf<uint32_t> src;     // sorted contributor -> index into the value array
};

struct Csr {
  Ctx* ctx = nullptr;
  int64_t nrows = 0, ncols = 0, nnz = 0;
  DBuf<int64_t> row_ptr;
  DBuf<int32_t> col;
  DBuf<double> val;
  // host copy behind feec_csr_view(..., on_device = 0), made on first request
  std::vector<int64_t> h_row_ptr;
  std::vector<int32_t> h_col;
  std::vector<double> h_val;
  bool have_host = false;
};

struct Mesh {
  Ctx* ctx = nullptr;
  int d = 0;
  int64_t n[5] = {0, 0, 0, 0, 0};
  // faces[k]: [ncells][C(d+1,k+1)] global index of each local face, k < d;
  // faces[0] are the (sorted) cell vertices.
  DBuf<int32_t> faces[5];
  // rep[k]: one (cell * nf + f) slot per k-simplex, 0 < k < d, decoding its vertices.
  DBuf<int64_t> rep[5];
  DBuf<uint8_t> bnd[5];
  // interior renumbering for the relative complex: -1 on the boundary
  DBuf<int32_t> interior[5];
  int64_t n_interior[5] = {0, 0, 0, 0, 0};
  bool have_interior = false;
  std::vector<int8_t> orientation;  // host, lazily
  bool have_orientation = false;
  bool orientable = false;
  int64_t components = -1;  // connected components (vertices joined by cells), lazily
  int kuhn_n = -1;          // cells per axis when a Kuhn grid (0: not one), lazily
  // cached assembly plans keyed by (bc, row grade, col grade)
  std::map<std::tuple<int, int, int>, std::unique_ptr<CooPlan>> plans;
  std::mutex mu;
  // exact topology on first request: Betti numbers by (bc, grade), integral
  // generators by (bc, grade, 0 cocycles / 1 cycles); guarded by topo_mu
  std::map<std::pair<int, int>, int64_t> betti;
  std::map<std::tuple<int, int, int>, std::vector<std::vector<int64_t>>> generators;
  std::mutex topo_mu;

  int nfaces(int k) const { return binom(d + 1, k + 1); }
  void ensure_interior();
  void ensure_orientation();
  void ensure_components();
};

struct Geom {
  Mesh* mesh = nullptr;
  DBuf<double> lengths;  // N_1 signed squared edge lengths
  // vertex coordinates [N_0][d] when the geometry came from coordinates (host
  // copy; the auxiliary-space preconditioner reads its frame off them)
  std::vector<double> coords;
  DBuf<double> dcoords;  // device copy of coords, made on first use (under mesh->mu)
  bool euclidean = false;
};

// ---------------------------------------------------------------------------
// module entry points (internal)
// ---------------------------------------------------------------------------
// topology.hip
void mesh_build(Mesh& m, int d, int64_t nverts, int64_t ncells, const int64_t* cells);
void mesh_coboundary(Mesh& m, int k, Csr& out);
void lengths_from_coords(Mesh& m, const double* coords_dev, int minkowski, double* out_dev);

// assemble.hip
void coo_plan_build(Ctx* ctx, CooPlan& plan, int64_t nrows, int64_t ncols, int64_t n,
                    const int32_t* rows, const int32_t* cols, const uint32_t* srcs);
// Segmented sum of vals[src] per plan entry; zero_drop removes entries whose
// contributors are all exactly 0.0 (the reference's `val != 0.0` push rule).
void coo_plan_apply(Ctx* ctx, const CooPlan& plan, const double* vals, bool zero_drop, Csr& out);
// Sparse utilities on device CSRs (nalgebra-sparse semantics).
void csr_transpose(const Csr& a, Csr& out);
void exclusive_scan_counts(Ctx* ctx, const unsigned long long* in, int64_t* out, int64_t n);
void csr_add(const Csr& a, const Csr& b, Csr& out);
void csr_mul(const Csr& a, const Csr& b, Csr& out);
void csr_select(const Csr& a, const int32_t* row_map, int64_t nrows, const int32_t* col_map,
                int64_t ncols, Csr& out);
struct BlockRef {
  const Csr* a;
  double scale;
  bool transpose;
};
void csr_block(Ctx* ctx, int nb, const BlockRef* blocks, const int64_t* row_sizes,
               const int64_t* col_sizes, Csr& out);
void csr_diagonal(const Csr& a, double* diag_dev);
int64_t coo_plan_values(Ctx* ctx, const CooPlan& plan, const double* vals, double* out,
                        uint8_t* keep);
struct CompactMap {
  DBuf<uint8_t> keep;
  DBuf<int64_t> pos;
  bool valid = false;
};
bool assemble_compacted(Ctx* ctx, const CooPlan& plan, const double* vals, CompactMap& cm,
                        Csr& A);
// Values of a plan whose columns are all 0 scattered to a dense vector of
// plan.nrows entries (rows without contributions get 0).
void coo_plan_dense_column(Ctx* ctx, const CooPlan& plan, const double* vals, double* out);
void form_contributions(Mesh& m, int bc, int rg, int cg, int r, int c, int64_t roff, int64_t coff,
                        bool transpose, uint32_t base, bool negate, int32_t* rows, int32_t* cols,
                        uint32_t* srcs);
int64_t ndofs(Mesh& m, int bc, int k);

// analytic.hip
// The geometry's vertex coordinates on the device, uploaded on first use
// (takes the mesh's lock; a geometry without coordinates is an argument error).
const double* device_coords(Geom& g);

// tune.hip
// Launches candidate `choice` of a bit-neutral launch choice through `launch`,
// the choice taken from the per-device table, or found by timing every
// candidate on the first call for (family, d, k, size class of n); returns it.
// timed = false launches candidate 0 unless a mode forces another.
int tune_run(Ctx* ctx, const char* family, int d, int k, int64_t n, int ncand,
             const std::function<void(int)>& launch, bool timed = true);

// elmat.hip
// Element matrices of one pairing, SoA [entry][cell].
void element_matrices_soa(Geom& g, int kind, int grade, double* out_dev);
int element_rows(int d, int kind, int grade);
int element_cols(int d, int kind, int grade);
void cell_metrics(Geom& g, double* gmat, double* ginv, double* det);
void kkt_element_blocks(Geom& g, int grade, double* const* outs, bool check);
struct BatchElementRequest {
  Geom* g;
  int grade;
  double* outs[7];
};
// All requests in one persistent launch; returns the kernel's seconds.
double kkt_element_blocks_batched(Ctx* ctx, const std::vector<BatchElementRequest>& reqs);
int dirac_element_count(int d);
void dirac_element_matrices(Geom& g, double* out_dev, bool check);
void causal_census(Geom& g, int64_t counts[3], uint8_t* riem_dev);

// assembly driver
void assemble_form(Geom& g, int bc, int kind, int grade, Csr& out);

// homology.hip: exact (co)homology on the host
// A sparse integer matrix by rows, each sorted by column with no zeros.
struct IntMat {
  int64_t nrows = 0, ncols = 0;
  std::vector<std::vector<std::pair<int32_t, int64_t>>> rows;
};
// the most harmonic forms the KKT border holds (the MINRES border kernels' limit)
constexpr int64_t kMaxHarmonics = 8;
IntMat int_transpose(const IntMat& a);
int64_t exact_rank(const IntMat& a);
std::vector<std::vector<int64_t>> quotient_generators(const IntMat& A, const IntMat& B);
int64_t betti_number(Mesh& m, int bc, int k);
std::vector<std::vector<int64_t>> cohomology_generators(Mesh& m, int bc, int k);
std::vector<std::vector<int64_t>> homology_generators(Mesh& m, int bc, int k);
void harmonics_exact(Geom& g, int bc, int k, std::vector<double>& integral, std::vector<double>& orthonormal,
                     int64_t& count);

// krylov.hip
struct KrylovStats {
  int64_t iters = 0;
  double precond_residual = 1.0;
  bool converged = false;
};
struct SolverState;  // opaque per-problem Krylov workspace
struct KktSystem {
  Ctx* ctx = nullptr;
  int64_t n = 0;        // KKT size without the harmonic border
  int64_t nh = 0;       // harmonic border columns
  Csr A;                // FP64 KKT (sigma,u blocks)
  DBuf<float> val32;    // FP32 copy of A's values
  DBuf<double> inv_diag;
  DBuf<float> inv_diag32;
  DBuf<double> border;  // [nh][n] dense border columns (M_k H) placed at u rows
  // Block preconditioner matrix blockdiag(hdif_gram(k-1), hdif_gram(k)), n x n
  // (nnz == 0: Jacobi with inv_diag), its FP32 values and the Chebyshev
  // coefficients of its two blocks (rows < split: block 0).
  Csr P;
  DBuf<int16_t> Pcol16;  // P's columns as col - row, when every |col - row| <= 32767
  bool band_checked = false;
  DBuf<float> P32;
  DBuf<_Float16> P16;    // D^-1/2 P D^-1/2 in FP16, CSR (when the SELL copy cannot be made)
  DBuf<float> sdinv32;   // D^-1/2, n entries
  // SELL-64 copy of D^-1/2 P D^-1/2 for the mixed path (made when the band fits
  // int16): rows in slices of 64, one wavefront lane per row; a slice's entries
  // are stored column-major (entry j of its 64 rows side by side), each packed
  // as {FP16 value | int16 (col' - row') << 16}, rows zero-padded to the slice's
  // longest; only off-diagonal entries, the unit diagonal is implied. Rows and
  // columns are renumbered (row' = sell_inv[row]) by sorting the rows of each
  // window of 256 (windows restart at `split`) by off-diagonal count, longest
  // first, so that a slice holds rows of equal length; sell_perm[row'] = row.
  // When some renumbered offset exceeds +-32767 (sell_wide), the slices keep
  // int32 columns in Psell_col instead and FP16 values in Psell_v16.
  DBuf<uint32_t> Psell;
  DBuf<int64_t> Psell_ptr;  // n_slices + 1 offsets into the slices (multiples of 64)
  DBuf<int32_t> sell_perm, sell_inv;
  int64_t sell_entries = 0, P_offdiag = 0;
  bool sell_wide = false;
  DBuf<int32_t> Psell_col;
  DBuf<_Float16> Psell_v16;
  // the same slices for the FP64 path: FP64 values, and int16 offsets apart
  // (or Psell_col when wide)
  DBuf<double> Psell64;
  DBuf<int16_t> Psell_off;
  DBuf<double> sdinv64;  // D^-1/2, n entries
  int64_t split = 0;
  int cheb_m = 0;
  DBuf<double> cheb;
  double lambda_max[2] = {0.0, 0.0};
  // multigrid / auxiliary-space block preconditioner (owned by the problem);
  // when set, solves use it instead of the Chebyshev polynomial
  struct Hx* hx = nullptr;
};
// Estimate the spectra of D^-1 P per block and set degree-m Chebyshev
// coefficients (m = 0 disables the polynomial).
void kkt_set_chebyshev(KktSystem& s, int m, double lower = 0.0);
void kkt_refresh_fp32(KktSystem& s);
// FP64 SELL-64 copy of the scaled preconditioner from P's values (when the
// pattern allows one).
void kkt_refresh_sell64(KktSystem& s);
// Solve s x = b (b, x device, FP64) to the options' tolerance.
feec_report kkt_solve(KktSystem& s, const double* b_dev, double* x_dev, const feec_solve_opts& o,
                      SolverState** ws);
void solver_state_free(SolverState* ws);
// Several solves whose MINRES runs advance in lockstep on lead's stream (one
// launch per phase for all of them); each result is bit-identical to
// kkt_solve's. Solves the lockstep path cannot take run one by one.
void kkt_solve_batch(std::vector<KktSystem*>& systems, std::vector<const double*>& b,
                     std::vector<double*>& x, const std::vector<feec_solve_opts>& opts,
                     std::vector<SolverState**>& ws, std::vector<feec_report>& reps, Ctx* lead);
void spmv64(Ctx* ctx, const Csr& a, const double* x, double* y);
// y = A x with one lane per row summing its entries in column order from the
// first product, FP contraction off: nalgebra-sparse's CSR x dense order.
void spmv_rowseq(Ctx* ctx, const Csr& a, const double* x, double* y);
double time_minres_spmv(KktSystem& s, bool fp32, int reps);
// z = B r with the block preconditioner as a solve of `precision` with
// `storage` (feec_pc_storage) applies it; device vectors of n + nh doubles.
void kkt_apply_preconditioner(KktSystem& s, int precision, int storage, const double* r, double* z);
// The feec_pc_format a solve of `precision` with `storage` applies.
int kkt_preconditioner_format(const KktSystem& s, int precision, int storage);
double time_cheb_step(KktSystem& s, bool fp32, int reps, int storage);
// Jacobi-preconditioned CG (krylov.rs:48-95: x0 = 0, stop when |r| <= rtol |b|)
// with its scalars on the device (host read every 8 iterations).
// The spectrum estimates' start vector (n entries on the device): 0.5 + U[0,1)
// from xorshift64 (13, 7, 17) seeded 0x2545F4914F6CDD1D, the i-th entry from
// the generator's (i+1)-th state.
void xorshift_start(Ctx* ctx, double* out, int64_t n);
KrylovStats cg_solve_device(Ctx* ctx, const Csr& a, const double* dinv, const double* b, double* x,
                            double rtol, int64_t max_iters);
// Sets up (once per device, that device current) the address of the scratch
// array of feec_csr_cg_noalloc on the current device.
void noalloc_scratch_init(int device);
// multigrid.hip
struct Hx;
struct Enq;
struct MinresS;
// n when the mesh is the Kuhn triangulation of the n^d box with the vertex
// numbering of feec_kuhn_grid (sum_a c_a (n+1)^a); 0 otherwise.
int kuhn_grid_n(Mesh& m);
// The HX block preconditioner diag(B_{k-1}, B_k, I) of a KKT system of grade k
// (values from s's preconditioner blocks), each block accelerated by a
// Chebyshev polynomial in B P with cheb_m applications of B (1: plain B) on
// [frac b, b] (lambda_max estimated with the FP32 cycle when spectrum_fp32,
// its launches replayed from a graph when use_graph);
// nullptr, with the reason, when it does not apply (no Kuhn
// grid, no coordinates, non-Euclidean, k > 2).
Hx* hx_create(Geom& g, int bc, int k, const KktSystem& s, int sweeps, int cheb_m, double frac,
              bool spectrum_fp32, bool use_graph, std::string* why);
void hx_info(const Hx& h, int* cheb_m, double lam[2], int* column_bytes);
// Seconds per launch: finest-level sweep, block preconditioner application
// (FP64_PC32), Chebyshev product by P, one B application (FP32).
void hx_kernel_times(Hx& h, int reps, bool use_graph, double out[4]);

// Host numerical helpers of the multigrid set-up.
// a (n x n, row-major, SPD) is replaced by its inverse, by Cholesky; false if
// a pivot is not positive.
bool spd_inverse(int64_t n, std::vector<double>& a);
// Largest eigenvalue of the symmetric tridiagonal matrix with diagonal diag
// and off-diagonal off (size n-1), by Sturm-count bisection.
double tridiagonal_lmax(const std::vector<double>& diag, const std::vector<double>& off);
// New values after re-assembly (same pattern).
void hx_refresh(Hx& h, Geom& g, const KktSystem& s);
void hx_free(Hx* h);
int64_t hx_bytes_per_apply(const Hx& h);
int hx_launches_per_apply(const Hx& h);
// z = B r, or p(B P) B r per block with the Chebyshev acceleration, with
// partial sums of r.z into part (red_grid(N) blocks): FP64
// vectors with the cycle in FP32 (fp32) or FP64, or FP32 throughout.
void hx_enqueue(Enq& q, Hx& h, const double* r, double* z, const MinresS* S, double* part, bool fp32);
void hx_enqueue(Enq& q, Hx& h, const float* r, float* z, const MinresS* S, double* part);

// a . b (part: 2 * 4096 doubles of scratch) and x += s a, on device doubles
double dot64(Ctx* ctx, const double* a, const double* b, int64_t n, double* part);
void axpy64(Ctx* ctx, const double* a, int64_t n, double s, double* x);

}  // namespace feec

struct feec_ctx_s : feec::Ctx {};
struct feec_mesh_s : feec::Mesh {};
struct feec_geom_s : feec::Geom {};
struct feec_csr_s : feec::Csr {};
