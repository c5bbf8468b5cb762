// The C ABI of libfeec (include/feec.h): handles, host-side meshers and
// synthetic inputs, topology and assembly queries, and the problem entry
// points built on the device kernels of the other translation units.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <vector>

#include "feec_internal.hpp"

namespace feec {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

// Device memory: a caching allocator over the device's default stream-ordered
// pool. A freed block goes to its context's cache, by size class, and is
// reused by that context's later allocations, in its stream's order. When a
// context is destroyed (its stream drained) its blocks move to a device-wide
// cache that any context may draw from. So a request's set-up and tear-down
// make no HIP call per buffer: hipFreeAsync costs about 17 us and serializes
// across threads, and a plain hipFree synchronizes the whole device. Blocks
// are returned to the pool only when an allocation fails for lack of memory.
namespace {
std::mutex g_cache_mu;
std::map<int, std::map<size_t, std::vector<void*>>> g_cache;  // device -> class -> blocks

// Size classes: multiples of 512 B up to 64 KiB, then 8 steps per power of two.
size_t size_class(size_t b) {
  if (b <= 65536) return (b + 511) & ~size_t(511);
  int e = 63 - __builtin_clzll((unsigned long long)(b - 1));  // 2^e < b <= 2^(e+1)
  const size_t step = (size_t)1 << (e - 2);
  return (b + step - 1) / step * step;
}

void* pop(std::map<size_t, std::vector<void*>>& c, size_t cls) {
  auto it = c.find(cls);
  if (it == c.end() || it->second.empty()) return nullptr;
  void* p = it->second.back();
  it->second.pop_back();
  return p;
}

// Give the cached blocks back to the pool: the device-wide ones (idle) and
// this context's own (freed in its stream's order).
void trim_caches(Ctx* ctx) {
  std::lock_guard<std::mutex> lock(g_cache_mu);
  for (auto* c : {&g_cache[ctx->device], &ctx->cache}) {
    for (auto& kv : *c)
      for (void* p : kv.second) (void)hipFreeAsync(p, ctx->stream);
    c->clear();
  }
  (void)hipStreamSynchronize(ctx->stream);
}
}  // namespace

void* device_alloc(Ctx* ctx, size_t bytes) {
  const size_t cls = size_class(std::max<size_t>(bytes, 1));
  void* p = pop(ctx->cache, cls);
  if (!p) {
    std::lock_guard<std::mutex> lock(g_cache_mu);
    p = pop(g_cache[ctx->device], cls);
  }
  if (!p) {
    hipError_t e = hipMallocAsync(&p, cls, ctx->stream);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      trim_caches(ctx);
      e = hipMallocAsync(&p, cls, ctx->stream);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fail(e == hipErrorOutOfMemory ? FEEC_E_OOM : FEEC_E_DEVICE,
           "hipMallocAsync(" + std::to_string(bytes) + " B): " + hipGetErrorString(e));
    }
  }
  ctx->bytes += (int64_t)cls;
  ++ctx->allocs;
  return p;
}

void device_free(Ctx* ctx, void* p, size_t bytes) {
  const size_t cls = size_class(std::max<size_t>(bytes, 1));
  ctx->cache[cls].push_back(p);
  ctx->bytes -= (int64_t)cls;
}

void release_cache(Ctx* ctx) {
  std::lock_guard<std::mutex> lock(g_cache_mu);
  auto& g = g_cache[ctx->device];
  for (auto& kv : ctx->cache) {
    auto& v = g[kv.first];
    v.insert(v.end(), kv.second.begin(), kv.second.end());
  }
  ctx->cache.clear();
}

void copy_to_device(Ctx* c, void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  FEEC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  c->sync();
}

void copy_to_host(Ctx* c, void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  FEEC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  c->sync();
}

void* Ctx::ensure_temp(size_t b) {
  if (b <= temp_bytes && temp) return temp;
  if (temp) {
    FEEC_HIP(hipStreamSynchronize(stream));
    device_free(this, temp, temp_bytes);
  }
  temp_bytes = std::max<size_t>(b, 1 << 20);
  temp = device_alloc(this, temp_bytes);
  return temp;
}

// SplitMix64 (Steele, Lea, Flood 2014): the generator of every synthetic input.
struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
};

// ---------------------------------------------------------------------------
// Mixed Hodge-Laplace KKT problem (elliptic.rs:54-182), device resident.
// ---------------------------------------------------------------------------
struct Problem {
  Geom* g = nullptr;
  int bc = 0, k = 0;
  feec_solve_opts opts{};
  int64_t ns = 0, nu = 0, nh = 0;
  int nfm = 0, nf = 0;
  // per-cell element blocks, one allocation: mass_sigma | dif_test | dif_both | diag_lo | diag_mk
  DBuf<double> el;
  int64_t off[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  CooPlan kkt_plan, diag_plan, pc_plan;
  CompactMap kkt_map;
  DBuf<double> diag;  // 2 (ns + nu)
  KktSystem sys;
  SolverState* ws = nullptr;
  DBuf<double> rhs, x, source;
  DBuf<double> harmonics;  // [nh][nu] M-orthonormal basis
  std::vector<double> harmonics_host;
  // b = M_k c for cochain sources: the full-complex M_k assembled from the M_k
  // element block (plan made on first use, values refreshed when the element
  // blocks change)
  CooPlan mass_plan;
  CompactMap mass_map;
  Csr Mk;
  bool have_mass_plan = false;
  int64_t el_version = 0, mass_version = -1;
  DBuf<double> load;
  // the multigrid / auxiliary-space preconditioner, when chosen
  Hx* hx = nullptr;
  bool hx_tried = false;
  feec_report last{};
  ~Problem() {
    solver_state_free(ws);
    hx_free(hx);
  }
};

namespace {

constexpr int TPB = 256;
inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + TPB - 1) / TPB); }

__global__ void k_inv_diag(const double* diag, int64_t n, int64_t nh, double* inv) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) inv[i] = 1.0 / (diag[i] + diag[n + i]);
  else if (i < n + nh) inv[i] = 1.0;
}

__global__ void k_rhs(const double* src, const int32_t* map, int64_t nfull, int64_t ns, int64_t N,
                      double* rhs) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nfull) return;
  int32_t j = map ? map[i] : (int32_t)i;
  if (j >= 0) rhs[ns + j] = src[i];
}

__global__ void k_extend(const double* x, const int32_t* map, int64_t nfull, int64_t off,
                         double* out) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nfull) return;
  int32_t j = map ? map[i] : (int32_t)i;
  out[i] = (j >= 0) ? x[off + j] : 0.0;
}

__global__ void k_soa_to_aos(const double* soa, int64_t nc, int per, double* aos) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nc * per) return;
  int64_t cell = t / per;
  int e = (int)(t % per);
  aos[t] = soa[(int64_t)e * nc + cell];
}

// vector contributions: entry rows[t] = base_row + map(faces[cell][i]), col 0
// The preconditioner's plan from the KKT plan. blockdiag(hdif_gram(k-1),
// hdif_gram(k)) has the pattern of the KKT's two diagonal blocks, whose plan
// entries hold one contribution per (cell, local pair): -M_{k-1} on the sigma
// block, dif_both(k+1) on the u block. The sort of P's own contributions (M
// pushed before dif_both) would give each entry the M contributions of the
// same (cell, pair)s in the same order, then the dif_both ones: these kernels
// write that plan without sorting. Per row: the entries of the row's own block
// (a contiguous run of the KKT row) and where the run starts.
__global__ void k_pc_rows(const int64_t* krp, const int32_t* kcol, int64_t n, int64_t ns,
                          unsigned long long* cnt, int64_t* kfirst) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r > n) return;
  if (r == n) {
    cnt[n] = 0;
    return;
  }
  const bool sig = r < ns;
  int64_t c = 0, first = -1;
  for (int64_t p = krp[r]; p < krp[r + 1]; ++p)
    if ((kcol[p] < ns) == sig) {
      if (first < 0) first = p;
      ++c;
    }
  cnt[r] = (unsigned long long)c;
  kfirst[r] = first < 0 ? krp[r] : first;
}

// each P entry's column, contribution count (twice the KKT entry's) and KKT entry
__global__ void k_pc_entries(const int64_t* prp, const int64_t* kfirst, const int32_t* kcol,
                             const int64_t* kseg, int64_t n, int32_t* pcol, unsigned long long* plen,
                             int64_t* pe) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int64_t t = 0; t < prp[r + 1] - prp[r]; ++t) {
    const int64_t e = kfirst[r] + t, q = prp[r] + t;
    pcol[q] = kcol[e];
    plen[q] = 2ull * (unsigned long long)(kseg[e + 1] - kseg[e]);
    pe[q] = e;
  }
}

// each P entry's sources (one lane per entry): the mass planes, then the
// dif_both planes, of the KKT entry's (cell, pair)s. An entry lies in the
// sigma block when its column does.
__global__ void k_pc_srcs(const int64_t* pe, const int32_t* pcol, const int64_t* kseg, const uint32_t* ksrc,
                          const int64_t* pseg, int64_t nnz, int64_t ns, uint32_t off_mlo, uint32_t off_dlo,
                          uint32_t off_mhi, uint32_t off_dhi, uint32_t* psrc) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= nnz) return;
  const bool sig = pcol[q] < ns;
  const int64_t e = pe[q];
  const int64_t L = kseg[e + 1] - kseg[e], base = pseg[q];
  for (int64_t i = 0; i < L; ++i) {
    const uint32_t s = ksrc[kseg[e] + i] & ~kCooNegate;
    const uint32_t local = s - (sig ? off_mlo : off_dhi);  // (pair * ncells + cell)
    psrc[base + i] = local + (sig ? off_mlo : off_mhi);
    psrc[base + L + i] = local + (sig ? off_dlo : off_dhi);
  }
}

__global__ void k_diag_coo(const int32_t* faces, int nf, const int32_t* map, int64_t nc, int64_t roff,
                           int stride_diag /* element entry of (i,i) = i*stride_diag */,
                           uint32_t base, int32_t* rows, int32_t* cols, uint32_t* srcs) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nc * nf) return;
  int64_t cell = t / nf;
  int i = (int)(t % nf);
  int32_t g = faces[cell * nf + i];
  if (map) g = map[g];
  rows[t] = g >= 0 ? (int32_t)(g + roff) : -1;
  cols[t] = g >= 0 ? 0 : -1;
  srcs[t] = base + (uint32_t)((int64_t)i * stride_diag * nc + cell);
}

__global__ void k_fill_border(const double* mh, int64_t nu, int64_t ns, int64_t n, int nh,
                              double* border) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int h = 0; h < nh; ++h) border[h * n + i] = (i >= ns) ? mh[h * nu + (i - ns)] : 0.0;
}

}  // namespace

// Harmonic basis (elliptic.rs:197-220), both readings [count][ndofs]: the
// period-normalized one (Harmonics::integral) and the M_k-orthonormal one the
// KKT border uses. Its dimension is b_k. On a Kuhn box (a ball) the closed
// forms: b_0 = 1 on the full complex, the constant; b_n = 1 on the relative
// complex, M_n^-1 times the orientation o (dirac.rs:553-568); every other
// space is trivial. On any other mesh b_k is computed exactly (modular ranks);
// the same closed forms are taken where b_k = 1 on a connected mesh, and
// otherwise the projected integral cocycles (harmonics_exact). A closed form
// is what harmonics_exact returns there: the cocycle projects to a multiple of
// 1 (resp. M_n^-1 o), whose period on the cycle (vertex 0, resp. o with its
// first cell positive) is made 1, then its M-norm.
void harmonic_bases(Geom& g, int bc, int k, std::vector<double>& orthonormal, std::vector<double>& integral,
                    int64_t& count) {
  Mesh& m = *g.mesh;
  FEEC_REQUIRE(k >= 0 && k <= m.d, FEEC_E_ARG, "grade out of range");
  count = 0;
  orthonormal.clear();
  integral.clear();
  const bool candidate = (bc == FEEC_FULL && k == 0) || (bc == FEEC_RELATIVE && k == m.d);
  bool closed = false;
  if (kuhn_grid_n(m) > 0) {
    if (!candidate) return;
    closed = true;
  } else {
    const int64_t b = betti_number(m, bc, k);
    if (b == 0) return;
    m.ensure_components();
    if (candidate && b == 1 && m.components == 1) {
      if (bc == FEEC_RELATIVE) m.ensure_orientation();
      closed = bc == FEEC_FULL || m.orientable;
    }
    if (!closed) {
      harmonics_exact(g, bc, k, integral, orthonormal, count);
      return;
    }
  }
  Csr M;
  assemble_form(g, bc, FEEC_MASS, k, M);
  int64_t n = M.nrows;
  std::vector<double> v(n);
  double period = 0.0;
  if (bc == FEEC_FULL) {
    std::fill(v.begin(), v.end(), 1.0);
    period = n ? 1.0 : 0.0;  // the cycle is vertex 0
  } else {
    m.ensure_orientation();
    FEEC_REQUIRE(m.orientable, FEEC_E_TOPOLOGY, "the mesh is not orientable");
    std::vector<int64_t> rp(n + 1);
    std::vector<int32_t> col(M.nnz);
    std::vector<double> val(M.nnz);
    copy_to_host(m.ctx, rp.data(), M.row_ptr.p, (n + 1) * 8);
    copy_to_host(m.ctx, col.data(), M.col.p, M.nnz * 4);
    copy_to_host(m.ctx, val.data(), M.val.p, M.nnz * 8);
    const double first = n && m.orientation[0] < 0 ? -1.0 : 1.0;
    for (int64_t i = 0; i < n; ++i) {
      double dgl = 0.0;
      for (int64_t p = rp[i]; p < rp[i + 1]; ++p)
        if (col[p] == i) dgl = val[p];
      const double o = first * (double)m.orientation[i];
      v[i] = o / dgl;  // top-grade Whitney mass is diagonal
      period += v[i] * o;
    }
  }
  FEEC_REQUIRE(period != 0.0, FEEC_E_DEVICE, "the harmonic form has no period");
  for (auto& x : v) x /= period;
  integral = v;
  // normalize in the M_k inner product
  DBuf<double> dv(m.ctx, n), mv(m.ctx, n);
  copy_to_device(m.ctx, dv.p, v.data(), n * 8);
  spmv64(m.ctx, M, dv.p, mv.p);
  m.ctx->sync();
  std::vector<double> hmv(n);
  copy_to_host(m.ctx, hmv.data(), mv.p, n * 8);
  double q = 0.0;
  for (int64_t i = 0; i < n; ++i) q += v[i] * hmv[i];
  FEEC_REQUIRE(q > 0.0, FEEC_E_UNSUPPORTED,
               "the L2 pairing is not definite on the harmonic space (the reference falls back to an "
               "eigensolve there, which is not built)");
  double s = 1.0 / std::sqrt(q);
  for (auto& x : v) x *= s;
  orthonormal = std::move(v);
  count = 1;
}

void harmonic_basis(Geom& g, int bc, int k, std::vector<double>& basis, int64_t& count) {
  std::vector<double> integral;
  harmonic_bases(g, bc, k, basis, integral, count);
}

void problem_build(Problem& P) {
  Geom& g = *P.g;
  Mesh& m = *g.mesh;
  Ctx* ctx = m.ctx;
  const int d = m.d, k = P.k, bc = P.bc;
  FEEC_REQUIRE(k >= 0 && k <= d, FEEC_E_ARG, "grade out of range");
  P.ns = ndofs(m, bc, k - 1);
  P.nu = ndofs(m, bc, k);
  P.nfm = (k >= 1) ? binom(d + 1, k) : 0;
  P.nf = binom(d + 1, k + 1);
  const int64_t nc = m.n[d];
  const int nfm = P.nfm, nf = P.nf;
  const bool has_up = k + 1 <= d;
  // element block layout
  // element blocks: 0 M_{k-1} | 1 dif_test(k) | 2 dif_both(k+1) | 3 diag dif_both(k) |
  // 4 diag M_k | 5 M_k | 6 dif_both(k)
  int64_t sizes[7] = {(int64_t)nfm * nfm * nc, (int64_t)nfm * nf * nc, has_up ? (int64_t)nf * nf * nc : 0,
                      (int64_t)nfm * nc, (int64_t)nf * nc, (int64_t)nf * nf * nc,
                      (int64_t)nfm * nfm * nc};
  P.off[0] = 0;
  for (int i = 0; i < 7; ++i) P.off[i + 1] = P.off[i] + sizes[i];
  FEEC_REQUIRE(P.off[7] < (int64_t(1) << 31), FEEC_E_ARG, "problem too large for 31-bit sources");
  P.el.alloc(ctx, std::max<int64_t>(P.off[7], 1));

  // harmonic border
  std::vector<double> hb;
  harmonic_basis(g, bc, k, hb, P.nh);
  P.harmonics_host = hb;

  // KKT contribution plan: [[-M_{k-1}, B], [B^T, dif_both(k+1)]]
  const int64_t n = P.ns + P.nu;
  int64_t total = sizes[0] + 2 * sizes[1] + sizes[2];
  {
    DBuf<int32_t> rows(ctx, std::max<int64_t>(total, 1)), cols(ctx, std::max<int64_t>(total, 1));
    DBuf<uint32_t> srcs(ctx, std::max<int64_t>(total, 1));
    int64_t at = 0;
    if (k >= 1) {
      form_contributions(m, bc, k - 1, k - 1, nfm, nfm, 0, 0, false, (uint32_t)P.off[0], true,
                         rows.p + at, cols.p + at, srcs.p + at);
      at += sizes[0];
      form_contributions(m, bc, k - 1, k, nfm, nf, 0, P.ns, false, (uint32_t)P.off[1], false,
                         rows.p + at, cols.p + at, srcs.p + at);
      at += sizes[1];
      form_contributions(m, bc, k - 1, k, nfm, nf, 0, P.ns, true, (uint32_t)P.off[1], false,
                         rows.p + at, cols.p + at, srcs.p + at);
      at += sizes[1];
    }
    if (has_up) {
      form_contributions(m, bc, k, k, nf, nf, P.ns, P.ns, false, (uint32_t)P.off[2], false,
                         rows.p + at, cols.p + at, srcs.p + at);
      at += sizes[2];
    }
    coo_plan_build(ctx, P.kkt_plan, n, n, at, rows.p, cols.p, srcs.p);
  }
  // diagonal plan: [diag M_{k-1} | diag M_k] then [diag dif_both(k) | diag dif_both(k+1)]
  {
    int64_t nd = nc * (2 * nfm + (has_up ? 2 : 1) * nf);
    DBuf<int32_t> rows(ctx, std::max<int64_t>(nd, 1)), cols(ctx, std::max<int64_t>(nd, 1));
    DBuf<uint32_t> srcs(ctx, std::max<int64_t>(nd, 1));
    if (bc == FEEC_RELATIVE) m.ensure_interior();
    const int32_t* map_lo = (bc == FEEC_RELATIVE && k >= 1) ? m.interior[k - 1].p : nullptr;
    const int32_t* map_k = (bc == FEEC_RELATIVE) ? m.interior[k].p : nullptr;
    int64_t at = 0;
    if (k >= 1) {
      k_diag_coo<<<grid_for(nc * nfm), TPB, 0, ctx->stream>>>(m.faces[k - 1].p, nfm, map_lo, nc, 0,
                                                               nfm + 1, (uint32_t)P.off[0],
                                                               rows.p + at, cols.p + at, srcs.p + at);
      at += nc * nfm;
      k_diag_coo<<<grid_for(nc * nfm), TPB, 0, ctx->stream>>>(m.faces[k - 1].p, nfm, map_lo, nc, n,
                                                               1, (uint32_t)P.off[3], rows.p + at,
                                                               cols.p + at, srcs.p + at);
      at += nc * nfm;
    }
    k_diag_coo<<<grid_for(nc * nf), TPB, 0, ctx->stream>>>(m.faces[k].p, nf, map_k, nc, P.ns, 1,
                                                            (uint32_t)P.off[4], rows.p + at,
                                                            cols.p + at, srcs.p + at);
    at += nc * nf;
    if (has_up) {
      k_diag_coo<<<grid_for(nc * nf), TPB, 0, ctx->stream>>>(m.faces[k].p, nf, map_k, nc, n + P.ns,
                                                              nf + 1, (uint32_t)P.off[2],
                                                              rows.p + at, cols.p + at, srcs.p + at);
      at += nc * nf;
    }
    FEEC_HIP(hipGetLastError());
    coo_plan_build(ctx, P.diag_plan, 2 * n, 1, at, rows.p, cols.p, srcs.p);
  }
  // preconditioner matrix blockdiag(hdif_gram(k-1), hdif_gram(k)) with
  // hdif_gram(j) = M_j + dif_both(j+1) (whitney_complex.rs:180-183)
  if (has_up) {
    // from the KKT plan (k_pc_rows): no sort of P's contributions
    CooPlan& pp = P.pc_plan;
    const CooPlan& kp = P.kkt_plan;
    pp.nrows = n;
    pp.ncols = n;
    pp.row_ptr.alloc(ctx, n + 1);
    DBuf<unsigned long long> cnt(ctx, n + 1);
    DBuf<int64_t> kfirst(ctx, std::max<int64_t>(n, 1));
    k_pc_rows<<<grid_for(n + 1), TPB, 0, ctx->stream>>>(kp.row_ptr.p, kp.col.p, n, P.ns, cnt.p, kfirst.p);
    FEEC_HIP(hipGetLastError());
    exclusive_scan_counts(ctx, cnt.p, pp.row_ptr.p, n + 1);
    copy_to_host(ctx, &pp.nnz, pp.row_ptr.p + n, 8);
    pp.col.alloc(ctx, pp.nnz);
    pp.seg.alloc(ctx, pp.nnz + 1);
    DBuf<unsigned long long> plen(ctx, pp.nnz + 1);
    DBuf<int64_t> pe(ctx, std::max<int64_t>(pp.nnz, 1));
    FEEC_HIP(hipMemsetAsync(plen.p + pp.nnz, 0, sizeof(unsigned long long), ctx->stream));
    if (n)
      k_pc_entries<<<grid_for(n), TPB, 0, ctx->stream>>>(pp.row_ptr.p, kfirst.p, kp.col.p, kp.seg.p, n, pp.col.p,
                                                         plen.p, pe.p);
    FEEC_HIP(hipGetLastError());
    exclusive_scan_counts(ctx, plen.p, pp.seg.p, pp.nnz + 1);
    int64_t nsrc = 0;
    copy_to_host(ctx, &nsrc, pp.seg.p + pp.nnz, 8);
    FEEC_REQUIRE(nsrc == sizes[0] + sizes[6] + sizes[5] + sizes[2] || bc == FEEC_RELATIVE, FEEC_E_DEVICE,
                 "preconditioner plan: contribution count");
    pp.src.alloc(ctx, std::max<int64_t>(nsrc, 1));
    if (pp.nnz)
      k_pc_srcs<<<grid_for(pp.nnz), TPB, 0, ctx->stream>>>(
          pe.p, pp.col.p, kp.seg.p, kp.src.p, pp.seg.p, pp.nnz, P.ns, (uint32_t)P.off[0], (uint32_t)P.off[6],
          (uint32_t)P.off[5], (uint32_t)P.off[2], pp.src.p);
    FEEC_HIP(hipGetLastError());
  } else {
    int64_t total_pc = sizes[0] + sizes[6] + sizes[5] + sizes[2];
    DBuf<int32_t> rows(ctx, std::max<int64_t>(total_pc, 1)), cols(ctx, std::max<int64_t>(total_pc, 1));
    DBuf<uint32_t> srcs(ctx, std::max<int64_t>(total_pc, 1));
    int64_t at = 0;
    if (k >= 1) {
      form_contributions(m, bc, k - 1, k - 1, nfm, nfm, 0, 0, false, (uint32_t)P.off[0], false,
                         rows.p + at, cols.p + at, srcs.p + at);
      at += sizes[0];
      form_contributions(m, bc, k - 1, k - 1, nfm, nfm, 0, 0, false, (uint32_t)P.off[6], false,
                         rows.p + at, cols.p + at, srcs.p + at);
      at += sizes[6];
    }
    form_contributions(m, bc, k, k, nf, nf, P.ns, P.ns, false, (uint32_t)P.off[5], false,
                       rows.p + at, cols.p + at, srcs.p + at);
    at += sizes[5];
    if (has_up) {
      form_contributions(m, bc, k, k, nf, nf, P.ns, P.ns, false, (uint32_t)P.off[2], false,
                         rows.p + at, cols.p + at, srcs.p + at);
      at += sizes[2];
    }
    coo_plan_build(ctx, P.pc_plan, n, n, at, rows.p, cols.p, srcs.p);
  }
  {
    Csr& Pm = P.sys.P;
    Pm.ctx = ctx;
    Pm.nrows = n;
    Pm.ncols = n;
    Pm.nnz = P.pc_plan.nnz;
    Pm.row_ptr.alloc(ctx, n + 1);
    Pm.col.alloc(ctx, std::max<int64_t>(Pm.nnz, 1));
    Pm.val.alloc(ctx, std::max<int64_t>(Pm.nnz, 1));
    FEEC_HIP(hipMemcpyAsync(Pm.row_ptr.p, P.pc_plan.row_ptr.p, (n + 1) * 8, hipMemcpyDeviceToDevice,
                            ctx->stream));
    if (Pm.nnz)
      FEEC_HIP(hipMemcpyAsync(Pm.col.p, P.pc_plan.col.p, Pm.nnz * 4, hipMemcpyDeviceToDevice,
                              ctx->stream));
    P.sys.split = P.ns;
  }
  P.diag.alloc(ctx, 2 * n);
  P.sys.ctx = ctx;
  P.sys.n = n;
  P.sys.nh = P.nh;
  P.sys.inv_diag.alloc(ctx, n + P.nh);
  P.rhs.alloc(ctx, n + P.nh);
  P.x.alloc(ctx, n + P.nh);
  P.source.alloc(ctx, std::max<int64_t>(m.n[k], 1));
  if (P.nh) {
    // border = M_k H, stored as full-length columns (zero on the sigma rows)
    Csr M;
    assemble_form(g, bc, FEEC_MASS, k, M);
    DBuf<double> h(ctx, P.nh * P.nu), mh(ctx, P.nh * P.nu);
    copy_to_device(ctx, h.p, P.harmonics_host.data(), P.nh * P.nu * 8);
    for (int64_t c = 0; c < P.nh; ++c) spmv64(ctx, M, h.p + c * P.nu, mh.p + c * P.nu);
    P.sys.border.alloc(ctx, P.nh * n);
    k_fill_border<<<grid_for(n), TPB, 0, ctx->stream>>>(mh.p, P.nu, P.ns, n, (int)P.nh,
                                                        P.sys.border.p);
    P.harmonics = std::move(h);
  }
  ctx->sync();
}

void problem_element_outputs(Problem& P, double** outs) {
  for (int i = 0; i < 7; ++i) outs[i] = P.el.p + P.off[i];
}

// cheb_degree < 0: by mesh resolution. A quasi-uniform mesh of N_0 vertices
// in d dimensions has about r = N_0^(1/d) vertices per axis and hdif_gram's
// condition number grows like r^2. The degree that balanced polynomial cost
// against Krylov iterations in sweeps on the mixed path (3D 64^3: 64; 2D 400^2
// and 800^2: 128; batches of 1e3-1e5-cell meshes: 24-32) is about 8 sqrt(r):
// taken in multiples of 8 within [16, 128], also for FP64 solves that run the
// polynomial in FP32. The FP64 path, whose steps move about 2.5x the bytes,
// takes half of it within [24, 64].
static int resolve_cheb_degree(const Problem& P) {
  if (P.opts.cheb_degree >= 0) return P.opts.cheb_degree;
  const Mesh& m = *P.g->mesh;
  const double r = std::pow((double)std::max<int64_t>(m.n[0], 1), 1.0 / m.d);
  int deg = 8 * (int)std::lround(std::sqrt(r));
  if (P.opts.precision == FEEC_PREC_FP64) return std::clamp(deg / 16 * 8, 24, 64);
  return std::clamp(deg, 16, 128);
}

// The multigrid / auxiliary-space preconditioner: built on the first
// assembly when the options ask for it (AUTO: where it applies and the KKT
// system has at least 20,000 rows), refreshed with the values afterwards.
constexpr int64_t kMultigridAutoRows = 20000;
// Chebyshev acceleration of the HX blocks: applications of B per block
// application (cheb_degree = -1) and the interval's lower end as a fraction of
// its upper one (cheb_lower <= 0).
constexpr int kMultigridChebDefault = 5;
constexpr double kMultigridChebFraction = 0.25;
void problem_multigrid(Problem& P) {
  const int pc = P.opts.preconditioner;
  if (pc == FEEC_PC_CHEBYSHEV || (pc == FEEC_PC_AUTO && P.sys.n < kMultigridAutoRows)) {
    P.sys.hx = nullptr;
    return;
  }
  if (P.hx) {
    hx_refresh(*P.hx, *P.g, P.sys);
    P.sys.hx = P.hx;
    return;
  }
  if (P.hx_tried && pc == FEEC_PC_AUTO) return;
  P.hx_tried = true;
  std::string why;
  const int m = P.opts.cheb_degree < 0 ? kMultigridChebDefault : std::max(1, P.opts.cheb_degree);
  const double frac = P.opts.cheb_lower > 0.0 ? P.opts.cheb_lower : kMultigridChebFraction;
  P.hx = hx_create(*P.g, P.bc, P.k, P.sys, P.opts.mg_sweeps, m, frac, P.opts.precision != FEEC_PREC_FP64,
                   P.opts.use_graph != 0, &why);
  FEEC_REQUIRE(P.hx || pc != FEEC_PC_MULTIGRID, FEEC_E_UNSUPPORTED,
               "multigrid preconditioner unavailable: " + why);
  P.sys.hx = P.hx;
  if (P.hx && P.ws) {  // graphs captured with the polynomial
    solver_state_free(P.ws);
    P.ws = nullptr;
  }
}

// CSR values, preconditioner diagonal and FP32 copy from the element blocks
// already in P.el.
void problem_assemble_from_elements(Problem& P) {
  Geom& g = *P.g;
  Mesh& m = *g.mesh;
  Ctx* ctx = m.ctx;
  ++P.el_version;
  bool rebuilt = assemble_compacted(ctx, P.kkt_plan, P.el.p, P.kkt_map, P.sys.A);
  coo_plan_dense_column(ctx, P.diag_plan, P.el.p, P.diag.p);
  k_inv_diag<<<grid_for(P.sys.n + P.nh), TPB, 0, ctx->stream>>>(P.diag.p, P.sys.n, P.nh,
                                                                P.sys.inv_diag.p);
  FEEC_HIP(hipGetLastError());
  if (rebuilt && P.ws) {  // a new pattern invalidates captured graphs
    solver_state_free(P.ws);
    P.ws = nullptr;
  }
  if (P.sys.P.nnz) coo_plan_values(ctx, P.pc_plan, P.el.p, P.sys.P.val.p, nullptr);
  problem_multigrid(P);
  // the FP32/FP16 copies are made for mixed and FP64_PC32 problems, the FP64
  // SELL copy for FP64 ones; either is refreshed once made
  if (P.opts.precision != FEEC_PREC_FP64 || P.sys.val32.p) kkt_refresh_fp32(P.sys);
  if (!P.hx) {
    if ((P.opts.precision == FEEC_PREC_FP64 && P.opts.precond_storage == FEEC_PC_STORE_SELL) ||
        P.sys.Psell64.p)
      kkt_refresh_sell64(P.sys);
    kkt_set_chebyshev(P.sys, resolve_cheb_degree(P), P.opts.cheb_lower);
  }
  ctx->sync();
}

void problem_assemble(Problem& P) {
  double* outs[7];
  problem_element_outputs(P, outs);
  kkt_element_blocks(*P.g, P.k, outs, true);
  problem_assemble_from_elements(P);
}

}  // namespace feec

using namespace feec;

struct feec_problem_s : feec::Problem {};

namespace feec {

// Rejects solver options outside their documented ranges.
static void validate_opts(const feec_solve_opts& o) {
  FEEC_REQUIRE(o.precision >= FEEC_PREC_FP64 && o.precision <= FEEC_PREC_FP64_PC32, FEEC_E_ARG,
               "precision must be a feec_precision value");
  FEEC_REQUIRE(o.precond_storage >= FEEC_PC_STORE_SELL && o.precond_storage <= FEEC_PC_STORE_CSR,
               FEEC_E_ARG, "precond_storage must be a feec_pc_storage value");
  FEEC_REQUIRE(o.precision != FEEC_PREC_FP64_PC32 || o.precond_storage == FEEC_PC_STORE_SELL,
               FEEC_E_ARG, "FEEC_PREC_FP64_PC32 runs on the SELL storage only");
  FEEC_REQUIRE(o.cheb_degree >= -1, FEEC_E_ARG, "cheb_degree must be >= -1");
  FEEC_REQUIRE(o.cheb_lower >= 0.0 && o.cheb_lower < 1.0, FEEC_E_ARG, "cheb_lower must be in [0, 1)");
  FEEC_REQUIRE(o.preconditioner >= FEEC_PC_AUTO && o.preconditioner <= FEEC_PC_MULTIGRID, FEEC_E_ARG,
               "preconditioner must be a feec_preconditioner value");
  FEEC_REQUIRE(o.mg_sweeps >= 1 && o.mg_sweeps <= 8, FEEC_E_ARG, "mg_sweeps must be 1..8");
}

Problem* problem_new(Geom* g, int bc, int k, const feec_solve_opts& o) {
  validate_opts(o);
  auto* P = new feec_problem_s();
  P->g = g;
  P->bc = bc;
  P->k = k;
  P->opts = o;
  try {
    problem_build(*P);
  } catch (...) {
    delete P;
    throw;
  }
  return P;
}

void problem_delete(Problem* P) { delete static_cast<feec_problem_s*>(P); }

int64_t problem_ndofs(Problem* P) { return P->sys.n + P->nh; }

// Upload a host source and form the KKT right-hand side (rhs = [0; E^T f; 0]).
void problem_prepare_rhs(Problem* P, const double* src) {
  Mesh& m = *P->g->mesh;
  Ctx* ctx = m.ctx;
  copy_to_device(ctx, P->source.p, src, m.n[P->k] * 8);
  const int64_t N = P->sys.n + P->nh;
  FEEC_HIP(hipMemsetAsync(P->rhs.p, 0, N * 8, ctx->stream));
  if (P->bc == FEEC_RELATIVE) m.ensure_interior();
  const int32_t* map = (P->bc == FEEC_RELATIVE) ? m.interior[P->k].p : nullptr;
  k_rhs<<<grid_for(m.n[P->k]), TPB, 0, ctx->stream>>>(P->source.p, map, m.n[P->k], P->ns, N,
                                                      P->rhs.p);
  FEEC_HIP(hipGetLastError());
}

// Upload a host source, solve; the report says whether it converged.
feec_report problem_solve_host(Problem* P, const double* src) {
  problem_prepare_rhs(P, src);
  feec_report r = kkt_solve(P->sys, P->rhs.p, P->x.p, P->opts, &P->ws);
  P->last = r;
  return r;
}

// Several problems (sources already uploaded by problem_prepare_rhs) solved
// with their MINRES runs in lockstep on lead's stream.
void problem_solve_batch(std::vector<Problem*>& probs, Ctx* lead, std::vector<feec_report>& reps) {
  std::vector<KktSystem*> sys;
  std::vector<const double*> b;
  std::vector<double*> x;
  std::vector<feec_solve_opts> o;
  std::vector<SolverState**> ws;
  for (Problem* P : probs) {
    sys.push_back(&P->sys);
    b.push_back(P->rhs.p);
    x.push_back(P->x.p);
    o.push_back(P->opts);
    ws.push_back(&P->ws);
  }
  kkt_solve_batch(sys, b, x, o, ws, reps, lead);
  for (size_t i = 0; i < probs.size(); ++i) probs[i]->last = reps[i];
}

void problem_set_report(Problem* P, const feec_report& r) { P->last = r; }

void problem_u(Problem* P, double* u) {
  Mesh& m = *P->g->mesh;
  Ctx* ctx = m.ctx;
  if (P->bc == FEEC_RELATIVE) m.ensure_interior();
  DBuf<double> o(ctx, m.n[P->k]);
  const int32_t* map = (P->bc == FEEC_RELATIVE) ? m.interior[P->k].p : nullptr;
  k_extend<<<grid_for(m.n[P->k]), TPB, 0, ctx->stream>>>(P->x.p, map, m.n[P->k], P->ns, o.p);
  ctx->sync();
  copy_to_host(ctx, u, o.p, m.n[P->k] * 8);
}

}  // namespace feec

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* feec_last_error(void) { return g_last_error.c_str(); }

const char* feec_status_string(feec_status s) {
  switch (abi_int(s)) {
    case FEEC_OK: return "ok";
    case FEEC_E_ARG: return "invalid argument";
    case FEEC_E_DEVICE: return "device error";
    case FEEC_E_OOM: return "out of device memory";
    case FEEC_E_TOPOLOGY: return "invalid topology";
    case FEEC_E_DEGENERATE_CELL: return "degenerate cell";
    case FEEC_E_NOT_SPD: return "not positive definite";
    case FEEC_E_NOT_CONVERGED: return "not converged";
    case FEEC_E_UNSUPPORTED: return "unsupported";
    case FEEC_E_ADMISSION: return "admission refused";
    case FEEC_E_AGAIN: return "not ready";
  }
  return "unknown";
}

void feec_solve_opts_default(feec_solve_opts* o) {
  if (!o) return;
  o->rtol = 1e-10;
  o->max_iters = 200000;
  o->precision = FEEC_PREC_FP64_PC32;
  o->max_refinements = 12;
  o->inner_rtol = 1e-4;
  o->use_graph = 1;
  o->graph_iters = 32;
  o->cheb_degree = -1;
  o->cheb_lower = 0.0;
  o->precond_storage = FEEC_PC_STORE_SELL;
  o->preconditioner = FEEC_PC_AUTO;
  o->mg_sweeps = 2;
}

feec_status feec_ctx_create(int device, feec_ctx** out) {
  return guard([&] {
    FEEC_REQUIRE(out, FEEC_E_ARG, "null output");
    int count = 0;
    FEEC_HIP(hipGetDeviceCount(&count));
    FEEC_REQUIRE(device >= 0 && device < count, FEEC_E_ARG, "no such device");
    FEEC_HIP(hipSetDevice(device));
    auto* c = new feec_ctx_s();
    c->device = device;
    int pools = 0;
    FEEC_HIP(hipDeviceGetAttribute(&pools, hipDeviceAttributeMemoryPoolsSupported, device));
    FEEC_REQUIRE(pools, FEEC_E_DEVICE, "device has no stream-ordered memory pools");
    hipMemPool_t pool = nullptr;
    FEEC_HIP(hipDeviceGetDefaultMemPool(&pool, device));
    uint64_t keep = UINT64_MAX;  // freed blocks stay cached in the pool
    FEEC_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc(&c->pinned, Ctx::kPinnedBytes, hipHostMallocDefault);
    if (e != hipSuccess) {
      if (c->stream) (void)hipStreamDestroy(c->stream);
      delete c;
      fail(FEEC_E_DEVICE, std::string("context setup: ") + hipGetErrorString(e));
    }
    try {
      c->dflag = static_cast<int*>(device_alloc(c, sizeof(int)));
      c->dcount = static_cast<unsigned long long*>(device_alloc(c, sizeof(unsigned long long)));
      noalloc_scratch_init(device);
    } catch (...) {
      feec_ctx_destroy(c);
      throw;
    }
    *out = c;
  });
}

void feec_ctx_destroy(feec_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->temp) device_free(ctx, ctx->temp, ctx->temp_bytes);
  if (ctx->dflag) device_free(ctx, ctx->dflag, sizeof(int));
  if (ctx->dcount) device_free(ctx, ctx->dcount, sizeof(unsigned long long));
  (void)hipStreamSynchronize(ctx->stream);
  release_cache(ctx);  // idle blocks, usable from any stream
  (void)hipStreamDestroy(ctx->stream);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  delete ctx;
}

feec_status feec_ctx_synchronize(feec_ctx* ctx) {
  return guard([&] { ctx->sync(); });
}

feec_status feec_ctx_memory(const feec_ctx* ctx, int64_t* bytes) {
  return guard([&] { *bytes = ctx->bytes.load(); });
}

feec_status feec_ctx_allocations(const feec_ctx* ctx, int64_t* count) {
  return guard([&] {
    FEEC_REQUIRE(ctx && count, FEEC_E_ARG, "null argument");
    *count = ctx->allocs.load();
  });
}

// ---- meshers --------------------------------------------------------------
feec_status feec_kuhn_counts(int dim, int n, int64_t* ncells, int64_t* nverts) {
  return guard([&] {
    FEEC_REQUIRE(dim >= 1 && dim <= 4 && n >= 1, FEEC_E_ARG, "dim 1..4, n >= 1");
    int64_t c = 1, v = 1, f = 1;
    for (int a = 0; a < dim; ++a) {
      c *= n;
      v *= (n + 1);
      f *= (a + 1);
    }
    if (ncells) *ncells = c * f;
    if (nverts) *nverts = v;
  });
}

// CartesianTopology::cell_skeleton (simplicial/src/mesher/grid.rs:77-98) and
// CartesianGrid::vertex_pos (regge/src/mesher/cartesian.rs:158-168).
feec_status feec_kuhn_grid(int dim, int n, int64_t* cells, double* coords) {
  return guard([&] {
    int64_t nc = 0, nv = 0;
    FEEC_REQUIRE(feec_kuhn_counts(dim, n, &nc, &nv) == FEEC_OK, FEEC_E_ARG, "bad grid");
    int64_t stride[4];
    int64_t s = 1;
    for (int a = 0; a < dim; ++a) {
      stride[a] = s;
      s *= (n + 1);
    }
    if (coords)
      for (int64_t v = 0; v < nv; ++v) {
        int64_t r = v;
        for (int a = 0; a < dim; ++a) {
          coords[v * dim + a] = (double)(r % (n + 1)) / (double)n;
          r /= (n + 1);
        }
      }
    if (!cells) return;
    int64_t nbox = 1;
    for (int a = 0; a < dim; ++a) nbox *= n;
    std::vector<int> perm(dim);
    int64_t out = 0;
    for (int64_t b = 0; b < nbox; ++b) {
      int64_t r = b, origin = 0;
      for (int a = 0; a < dim; ++a) {
        origin += (r % n) * stride[a];
        r /= n;
      }
      std::iota(perm.begin(), perm.end(), 0);
      do {
        int64_t v = origin;
        cells[out * (dim + 1)] = v;
        for (int i = 0; i < dim; ++i) {
          v += stride[perm[i]];
          cells[out * (dim + 1) + i + 1] = v;
        }
        ++out;
      } while (std::next_permutation(perm.begin(), perm.end()));
    }
  });
}

feec_status feec_jitter_interior(int dim, int n, double* coords, double amplitude, uint64_t seed,
                                 uint32_t axis_mask) {
  return guard([&] {
    FEEC_REQUIRE(dim >= 1 && dim <= 4 && n >= 1 && coords, FEEC_E_ARG, "bad jitter arguments");
    int64_t nv = 1;
    for (int a = 0; a < dim; ++a) nv *= (n + 1);
    SplitMix64 rng{seed};
    const double h = 1.0 / n;
    for (int64_t v = 0; v < nv; ++v) {
      int64_t r = v;
      bool interior = true;
      for (int a = 0; a < dim; ++a) {
        int64_t c = r % (n + 1);
        r /= (n + 1);
        if (c == 0 || c == n) interior = false;
      }
      if (!interior) continue;
      for (int a = 0; a < dim; ++a)
        if (axis_mask & (1u << a)) coords[v * dim + a] += amplitude * h * (2.0 * rng.uniform() - 1.0);
    }
  });
}

feec_status feec_normal_vector(int64_t n, uint64_t seed, double* out) {
  return guard([&] {
    FEEC_REQUIRE(n >= 0 && (n == 0 || out), FEEC_E_ARG, "bad output");
    SplitMix64 rng{seed};
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < n; i += 2) {
      double u1 = 1.0 - rng.uniform();  // (0, 1]
      double u2 = rng.uniform();
      double rad = std::sqrt(-2.0 * std::log(u1));
      out[i] = rad * std::cos(two_pi * u2);
      if (i + 1 < n) out[i + 1] = rad * std::sin(two_pi * u2);
    }
  });
}

// ---- topology ---------------------------------------------------------------
feec_status feec_mesh_create(feec_ctx* ctx, int dim, int64_t nverts, int64_t ncells,
                             const int64_t* cells, feec_mesh** out) {
  return guard([&] {
    FEEC_REQUIRE(ctx && cells && out, FEEC_E_ARG, "null argument");
    FEEC_HIP(hipSetDevice(ctx->device));
    auto* m = new feec_mesh_s();
    m->ctx = ctx;
    try {
      mesh_build(*m, dim, nverts, ncells, cells);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

void feec_mesh_destroy(feec_mesh* m) { delete m; }

feec_status feec_mesh_dim(const feec_mesh* m, int* dim) {
  return guard([&] { *dim = m->d; });
}

feec_status feec_mesh_counts(const feec_mesh* m, int64_t* counts) {
  return guard([&] {
    for (int k = 0; k <= m->d; ++k) counts[k] = m->n[k];
  });
}

feec_status feec_mesh_skeleton(const feec_mesh* mc, int k, int64_t* verts) {
  return guard([&] {
    feec_mesh* m = const_cast<feec_mesh*>(mc);
    FEEC_REQUIRE(k >= 0 && k <= m->d && verts, FEEC_E_ARG, "bad grade");
    const int d = m->d, nv = d + 1;
    const int64_t nc = m->n[d];
    std::vector<int32_t> cells((size_t)nc * nv);
    copy_to_host(m->ctx, cells.data(), m->faces[0].p, cells.size() * 4);
    if (k == 0) {
      for (int64_t v = 0; v < m->n[0]; ++v) verts[v] = v;
      return;
    }
    if (k == d) {
      for (size_t i = 0; i < cells.size(); ++i) verts[i] = cells[i];
      return;
    }
    std::vector<int64_t> rep(m->n[k]);
    copy_to_host(m->ctx, rep.data(), m->rep[k].p, rep.size() * 8);
    std::vector<std::vector<int>> subs;
    for (unsigned msk = 0; msk < (1u << nv); ++msk) {
      if (__builtin_popcount(msk) != k + 1) continue;
      std::vector<int> e;
      for (int b = 0; b < nv; ++b)
        if (msk & (1u << b)) e.push_back(b);
      subs.push_back(e);
    }
    const int nf = (int)subs.size();
    for (int64_t s = 0; s < m->n[k]; ++s) {
      int64_t cell = rep[s] / nf;
      int f = (int)(rep[s] % nf);
      for (int p = 0; p <= k; ++p) verts[s * (k + 1) + p] = cells[cell * nv + subs[f][p]];
    }
  });
}

feec_status feec_mesh_cell_faces(const feec_mesh* m, int k, int64_t* out) {
  return guard([&] {
    FEEC_REQUIRE(k >= 0 && k <= m->d && out, FEEC_E_ARG, "bad grade");
    int64_t n = m->n[m->d] * (int64_t)binom(m->d + 1, k + 1);
    std::vector<int32_t> h(n);
    copy_to_host(m->ctx, h.data(), m->faces[k].p, n * 4);
    for (int64_t i = 0; i < n; ++i) out[i] = h[i];
  });
}

feec_status feec_mesh_boundary(const feec_mesh* m, int k, uint8_t* flags) {
  return guard([&] {
    FEEC_REQUIRE(k >= 0 && k <= m->d && flags, FEEC_E_ARG, "bad grade");
    copy_to_host(m->ctx, flags, m->bnd[k].p, m->n[k]);
  });
}

feec_status feec_mesh_orientation(feec_mesh* m, int8_t* signs) {
  return guard([&] {
    m->ensure_orientation();
    FEEC_REQUIRE(m->orientable, FEEC_E_TOPOLOGY, "the mesh is not orientable");
    std::copy(m->orientation.begin(), m->orientation.end(), signs);
  });
}

feec_status feec_mesh_coboundary(feec_mesh* m, feec_bc bc, int k, feec_csr** out) {
  return guard([&] {
    FEEC_REQUIRE(m && out, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(abi_bc_ok(bc), FEEC_E_ARG, "bc must be a feec_bc value");
    FEEC_HIP(hipSetDevice(m->ctx->device));
    auto* a = new feec_csr_s();
    a->ctx = m->ctx;
    try {
      if (k < 0 || k >= m->d) {  // HilbertComplex::dif is total in grade
        a->nrows = ndofs(*m, bc, k + 1);
        a->ncols = ndofs(*m, bc, k);
        a->nnz = 0;
        a->row_ptr.alloc(m->ctx, a->nrows + 1);
        FEEC_HIP(hipMemsetAsync(a->row_ptr.p, 0, (a->nrows + 1) * 8, m->ctx->stream));
        m->ctx->sync();
      } else if (bc == FEEC_RELATIVE) {
        Csr full;
        mesh_coboundary(*m, k, full);
        m->ensure_interior();
        csr_select(full, m->interior[k + 1].p, m->n_interior[k + 1], m->interior[k].p,
                   m->n_interior[k], *a);
      } else {
        mesh_coboundary(*m, k, *a);
      }
    } catch (...) {
      delete a;
      throw;
    }
    *out = a;
  });
}

// ---- geometry ---------------------------------------------------------------
feec_status feec_geom_from_lengths(feec_mesh* m, const double* lengths, feec_geom** out) {
  return guard([&] {
    FEEC_REQUIRE(m && lengths && out, FEEC_E_ARG, "null argument");
    FEEC_HIP(hipSetDevice(m->ctx->device));
    auto* g = new feec_geom_s();
    g->mesh = m;
    g->lengths.alloc(m->ctx, m->n[1]);
    copy_to_device(m->ctx, g->lengths.p, lengths, m->n[1] * 8);
    *out = g;
  });
}

feec_status feec_geom_from_coords(feec_mesh* m, const double* coords, feec_ambient amb,
                                  feec_geom** out) {
  return guard([&] {
    FEEC_REQUIRE(m && coords && out, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(abi_int(amb) == FEEC_EUCLIDEAN || abi_int(amb) == FEEC_MINKOWSKI, FEEC_E_ARG,
                 "ambient must be a feec_ambient value");
    FEEC_HIP(hipSetDevice(m->ctx->device));
    auto* g = new feec_geom_s();
    g->mesh = m;
    g->lengths.alloc(m->ctx, m->n[1]);
    DBuf<double> c(m->ctx, m->n[0] * m->d);
    copy_to_device(m->ctx, c.p, coords, m->n[0] * m->d * 8);
    lengths_from_coords(*m, c.p, amb == FEEC_MINKOWSKI, g->lengths.p);
    g->coords.assign(coords, coords + m->n[0] * m->d);
    g->euclidean = amb != FEEC_MINKOWSKI;
    m->ctx->sync();
    *out = g;
  });
}

void feec_geom_destroy(feec_geom* g) { delete g; }

feec_status feec_geom_lengths(const feec_geom* g, double* out) {
  return guard([&] {
    copy_to_host(g->mesh->ctx, out, g->lengths.p, g->mesh->n[1] * 8);
  });
}

feec_status feec_cell_metrics(feec_geom* g, double* gm, double* gi, double* det) {
  return guard([&] {
    Mesh& m = *g->mesh;
    int64_t nc = m.n[m.d], dd = (int64_t)m.d * m.d;
    DBuf<double> a(m.ctx, gm ? nc * dd : 1), b(m.ctx, gi ? nc * dd : 1), c(m.ctx, det ? nc : 1);
    cell_metrics(*g, gm ? a.p : nullptr, gi ? b.p : nullptr, det ? c.p : nullptr);
    m.ctx->sync();
    if (gm) copy_to_host(m.ctx, gm, a.p, nc * dd * 8);
    if (gi) copy_to_host(m.ctx, gi, b.p, nc * dd * 8);
    if (det) copy_to_host(m.ctx, det, c.p, nc * 8);
  });
}

feec_status feec_causal_census(feec_geom* g, int64_t counts[3], uint8_t* riem) {
  return guard([&] {
    Mesh& m = *g->mesh;
    DBuf<uint8_t> r(m.ctx, riem ? m.n[m.d] : 1);
    causal_census(*g, counts, riem ? r.p : nullptr);
    if (riem) copy_to_host(m.ctx, riem, r.p, m.n[m.d]);
  });
}

// ---- element matrices and assembly ---------------------------------------------
feec_status feec_element_shape(int dim, feec_form_kind kind, int grade, int* rows, int* cols) {
  return guard([&] {
    FEEC_REQUIRE(rows && cols, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(abi_int(kind) >= FEEC_MASS && abi_int(kind) <= FEEC_DIF_BOTH, FEEC_E_ARG,
                 "not an element form");
    *rows = (grade < 0 || grade > dim) ? 0 : element_rows(dim, kind, grade);
    *cols = (grade < 0 || grade > dim) ? 0 : element_cols(dim, kind, grade);
  });
}

feec_status feec_element_matrices(feec_geom* g, feec_form_kind kind, int grade, double* out) {
  return guard([&] {
    FEEC_REQUIRE(g && out, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(abi_int(kind) >= FEEC_MASS && abi_int(kind) <= FEEC_DIF_BOTH, FEEC_E_ARG,
                 "not an element form");
    Mesh& m = *g->mesh;
    FEEC_REQUIRE(grade >= 0 && grade <= m.d, FEEC_E_ARG, "grade out of range");
    int r = element_rows(m.d, kind, grade), c = element_cols(m.d, kind, grade);
    int64_t nc = m.n[m.d], total = nc * r * c;
    if (total == 0) return;
    DBuf<double> soa(m.ctx, total), aos(m.ctx, total);
    element_matrices_soa(*g, kind, grade, soa.p);
    k_soa_to_aos<<<grid_for(total), TPB, 0, m.ctx->stream>>>(soa.p, nc, r * c, aos.p);
    FEEC_HIP(hipGetLastError());
    m.ctx->sync();
    copy_to_host(m.ctx, out, aos.p, total * 8);
  });
}

feec_status feec_assemble(feec_geom* g, feec_bc bc, feec_form_kind kind, int grade,
                          feec_csr** out) {
  return guard([&] {
    FEEC_REQUIRE(g && out, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(abi_int(kind) >= FEEC_MASS && abi_int(kind) <= FEEC_HDIF_GRAM, FEEC_E_ARG,
                 "bad form kind");
    FEEC_REQUIRE(abi_bc_ok(bc), FEEC_E_ARG, "bc must be a feec_bc value");
    FEEC_HIP(hipSetDevice(g->mesh->ctx->device));
    auto* a = new feec_csr_s();
    try {
      assemble_form(*g, bc, kind, grade, *a);
    } catch (...) {
      delete a;
      throw;
    }
    *out = a;
  });
}

feec_status feec_csr_shape(const feec_csr* a, int64_t* nrows, int64_t* ncols, int64_t* nnz) {
  return guard([&] {
    if (nrows) *nrows = a->nrows;
    if (ncols) *ncols = a->ncols;
    if (nnz) *nnz = a->nnz;
  });
}

feec_status feec_csr_copy(const feec_csr* a, int64_t* row_ptr, int32_t* col, double* val) {
  return guard([&] {
    if (row_ptr) copy_to_host(a->ctx, row_ptr, a->row_ptr.p, (a->nrows + 1) * 8);
    if (col && a->nnz) copy_to_host(a->ctx, col, a->col.p, a->nnz * 4);
    if (val && a->nnz) copy_to_host(a->ctx, val, a->val.p, a->nnz * 8);
  });
}

feec_status feec_csr_spmv(feec_csr* a, const double* x, double* y) {
  return guard([&] {
    DBuf<double> dx(a->ctx, std::max<int64_t>(a->ncols, 1)), dy(a->ctx, std::max<int64_t>(a->nrows, 1));
    copy_to_device(a->ctx, dx.p, x, a->ncols * 8);
    spmv64(a->ctx, *a, dx.p, dy.p);
    a->ctx->sync();
    copy_to_host(a->ctx, y, dy.p, a->nrows * 8);
  });
}

feec_status feec_csr_view(const feec_csr* a, int64_t* nrows, int64_t* ncols, int64_t* nnz,
                          const int64_t** row_ptr, const int32_t** col, const double** val,
                          int on_device) {
  return guard([&] {
    FEEC_REQUIRE(a, FEEC_E_ARG, "null matrix");
    if (nrows) *nrows = a->nrows;
    if (ncols) *ncols = a->ncols;
    if (nnz) *nnz = a->nnz;
    if (on_device) {
      // the arrays are written on the context's stream: drain it first
      a->ctx->sync();
      if (row_ptr) *row_ptr = a->row_ptr.p;
      if (col) *col = a->col.p;
      if (val) *val = a->val.p;
      return;
    }
    auto* h = const_cast<feec_csr*>(a);
    if (!h->have_host) {
      h->h_row_ptr.resize(a->nrows + 1);
      h->h_col.resize(a->nnz);
      h->h_val.resize(a->nnz);
      copy_to_host(a->ctx, h->h_row_ptr.data(), a->row_ptr.p, (a->nrows + 1) * 8);
      if (a->nnz) {
        copy_to_host(a->ctx, h->h_col.data(), a->col.p, a->nnz * 4);
        copy_to_host(a->ctx, h->h_val.data(), a->val.p, a->nnz * 8);
      }
      h->have_host = true;
    }
    if (row_ptr) *row_ptr = h->h_row_ptr.data();
    if (col) *col = h->h_col.data();
    if (val) *val = h->h_val.data();
  });
}

feec_status feec_csr_spmv_device(const feec_csr* a, const double* x, double* y) {
  return guard([&] {
    FEEC_REQUIRE(a && (x || !a->ncols) && (y || !a->nrows), FEEC_E_ARG, "null argument");
    FEEC_HIP(hipSetDevice(a->ctx->device));
    spmv64(a->ctx, *a, x, y);
    a->ctx->sync();
  });
}

feec_status feec_vec_dot_device(feec_ctx* ctx, int64_t n, const double* x, const double* y,
                                double* out) {
  return guard([&] {
    FEEC_REQUIRE(ctx && out && n >= 0 && (n == 0 || (x && y)), FEEC_E_ARG, "bad argument");
    FEEC_HIP(hipSetDevice(ctx->device));
    if (n == 0) {
      *out = 0.0;
      return;
    }
    DBuf<double> part(ctx, 2 * 4096);
    *out = dot64(ctx, x, y, n, part.p);
  });
}

feec_status feec_vec_axpy_device(feec_ctx* ctx, int64_t n, double a, const double* x, double* y) {
  return guard([&] {
    FEEC_REQUIRE(ctx && n >= 0 && (n == 0 || (x && y)), FEEC_E_ARG, "bad argument");
    FEEC_HIP(hipSetDevice(ctx->device));
    axpy64(ctx, x, n, a, y);
    ctx->sync();
  });
}

feec_status feec_load_vector(feec_geom* g, int grade, const double* cochain, double* out) {
  return guard([&] {
    FEEC_REQUIRE(g && cochain && out, FEEC_E_ARG, "null argument");
    Mesh& m = *g->mesh;
    FEEC_REQUIRE(grade >= 0 && grade <= m.d, FEEC_E_ARG, "grade out of range");
    Ctx* ctx = m.ctx;
    FEEC_HIP(hipSetDevice(ctx->device));
    Csr M;
    assemble_form(*g, FEEC_FULL, FEEC_MASS, grade, M);
    const int64_t n = m.n[grade];
    DBuf<double> dc(ctx, std::max<int64_t>(n, 1)), db(ctx, std::max<int64_t>(n, 1));
    copy_to_device(ctx, dc.p, cochain, n * 8);
    spmv_rowseq(ctx, M, dc.p, db.p);
    copy_to_host(ctx, out, db.p, n * 8);
  });
}

void feec_csr_destroy(feec_csr* a) { delete a; }

namespace {
__global__ void k_inverse(const double* d, int64_t n, double* out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = d[i] != 0.0 ? 1.0 / d[i] : 1.0;
}
}  // namespace

feec_status feec_csr_cg(feec_csr* a, const double* b, double* x, double rtol, int64_t max_iters,
                        feec_report* report) {
  return guard([&] {
    FEEC_REQUIRE(a && b && x, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(a->nrows == a->ncols, FEEC_E_ARG, "CG needs a square matrix");
    FEEC_REQUIRE(rtol >= 0.0 && max_iters >= 0, FEEC_E_ARG, "rtol >= 0 and max_iters >= 0");
    Ctx* ctx = a->ctx;
    FEEC_HIP(hipSetDevice(ctx->device));
    const int64_t n = a->nrows;
    const int64_t nb = std::max<int64_t>(n, 1);
    DBuf<double> db(ctx, nb), dx(ctx, nb), dinv(ctx, nb);
    copy_to_device(ctx, db.p, b, n * 8);
    // Jacobi (precond.rs:53-123): the inverse diagonal, 1 where it is zero
    csr_diagonal(*a, dinv.p);
    k_inverse<<<(unsigned)std::max<int64_t>(1, (n + 255) / 256), 256, 0, ctx->stream>>>(dinv.p, n,
                                                                                      dinv.p);
    FEEC_HIP(hipGetLastError());
    KrylovStats ks = cg_solve_device(ctx, *a, dinv.p, db.p, dx.p, rtol, max_iters);
    copy_to_host(ctx, x, dx.p, n * 8);
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

// ---- problems ---------------------------------------------------------------------
feec_status feec_harmonics(feec_geom* g, feec_bc bc, int grade, int64_t* count, double* basis) {
  return guard([&] {
    FEEC_REQUIRE(g, FEEC_E_ARG, "null geometry");
    FEEC_REQUIRE(abi_bc_ok(bc), FEEC_E_ARG, "bc must be a feec_bc value");
    std::vector<double> hb;
    int64_t c = 0;
    harmonic_basis(*g, bc, grade, hb, c);
    *count = c;
    if (basis && c) std::copy(hb.begin(), hb.end(), basis);
  });
}

feec_status feec_harmonics_integral(feec_geom* g, feec_bc bc, int grade, int64_t* count, double* basis) {
  return guard([&] {
    FEEC_REQUIRE(g && count, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(abi_bc_ok(bc), FEEC_E_ARG, "bc must be a feec_bc value");
    std::vector<double> ortho, integral;
    int64_t c = 0;
    harmonic_bases(*g, bc, grade, ortho, integral, c);
    *count = c;
    if (basis && c) std::copy(integral.begin(), integral.end(), basis);
  });
}

feec_status feec_betti_number(const feec_mesh* mc, feec_bc bc, int grade, int64_t* b) {
  return guard([&] {
    FEEC_REQUIRE(mc && b, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(abi_bc_ok(bc), FEEC_E_ARG, "bc must be a feec_bc value");
    feec_mesh* m = const_cast<feec_mesh*>(mc);
    FEEC_HIP(hipSetDevice(m->ctx->device));
    *b = betti_number(*m, bc, grade);
  });
}

namespace {
feec_status generators_out(const feec_mesh* mc, feec_bc bc, int grade, int64_t* count, int64_t* out,
                           bool cocycles) {
  return guard([&] {
    FEEC_REQUIRE(mc && count, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(abi_bc_ok(bc), FEEC_E_ARG, "bc must be a feec_bc value");
    feec_mesh* m = const_cast<feec_mesh*>(mc);
    FEEC_HIP(hipSetDevice(m->ctx->device));
    const auto gens = cocycles ? cohomology_generators(*m, bc, grade) : homology_generators(*m, bc, grade);
    *count = (int64_t)gens.size();
    if (out)
      for (size_t i = 0; i < gens.size(); ++i) std::copy(gens[i].begin(), gens[i].end(), out + i * gens[i].size());
  });
}
}  // namespace

feec_status feec_cohomology_generators(const feec_mesh* m, feec_bc bc, int grade, int64_t* count, int64_t* out) {
  return generators_out(m, bc, grade, count, out, true);
}

feec_status feec_homology_generators(const feec_mesh* m, feec_bc bc, int grade, int64_t* count, int64_t* out) {
  return generators_out(m, bc, grade, count, out, false);
}

feec_status feec_mixed_kkt(feec_geom* g, feec_bc bc, int grade, feec_csr** out, int64_t sizes[3]) {
  return guard([&] {
    FEEC_REQUIRE(g && out, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(abi_bc_ok(bc), FEEC_E_ARG, "bc must be a feec_bc value");
    Problem P;
    P.g = g;
    P.bc = bc;
    P.k = grade;
    feec_solve_opts_default(&P.opts);
    problem_build(P);
    problem_assemble(P);
    Ctx* ctx = g->mesh->ctx;
    auto* a = new feec_csr_s();
    a->ctx = ctx;
    if (P.nh == 0) {
      a->nrows = P.sys.A.nrows;
      a->ncols = P.sys.A.ncols;
      a->nnz = P.sys.A.nnz;
      a->row_ptr = std::move(P.sys.A.row_ptr);
      a->col = std::move(P.sys.A.col);
      a->val = std::move(P.sys.A.val);
    } else {
      // the border as dense CSR blocks, every entry kept (elliptic.rs:62-84)
      int64_t n = P.sys.n, nh = P.nh, nu = P.nu;
      std::vector<double> bh(nh * n);
      copy_to_host(ctx, bh.data(), P.sys.border.p, nh * n * 8);
      Csr col_blk, row_blk;
      std::vector<int64_t> rp(nu + 1), rp2(nh + 1);
      std::vector<int32_t> c1(nu * nh), c2(nh * nu);
      std::vector<double> v1(nu * nh), v2(nh * nu);
      for (int64_t i = 0; i < nu; ++i) {
        rp[i] = i * nh;
        for (int64_t h = 0; h < nh; ++h) {
          c1[i * nh + h] = (int32_t)h;
          v1[i * nh + h] = bh[h * n + P.ns + i];
        }
      }
      rp[nu] = nu * nh;
      for (int64_t h = 0; h < nh; ++h) {
        rp2[h] = h * nu;
        for (int64_t i = 0; i < nu; ++i) {
          c2[h * nu + i] = (int32_t)i;
          v2[h * nu + i] = bh[h * n + P.ns + i];
        }
      }
      rp2[nh] = nh * nu;
      auto up = [&](Csr& c, int64_t r, int64_t cc, std::vector<int64_t>& rpv,
                    std::vector<int32_t>& cv, std::vector<double>& vv) {
        c.ctx = ctx;
        c.nrows = r;
        c.ncols = cc;
        c.nnz = (int64_t)cv.size();
        c.row_ptr.alloc(ctx, r + 1);
        c.col.alloc(ctx, c.nnz);
        c.val.alloc(ctx, c.nnz);
        copy_to_device(ctx, c.row_ptr.p, rpv.data(), (r + 1) * 8);
        copy_to_device(ctx, c.col.p, cv.data(), c.nnz * 4);
        copy_to_device(ctx, c.val.p, vv.data(), c.nnz * 8);
      };
      up(col_blk, nu, nh, rp, c1, v1);
      up(row_blk, nh, nu, rp2, c2, v2);
      // [[A_ss, A_su, 0],[A_us, A_uu, MH],[0, (MH)^T, 0]] from the KKT plus the border
      BlockRef blocks[4] = {{&P.sys.A, 1.0, false}, {nullptr, 1.0, false}, {nullptr, 1.0, false},
                            {nullptr, 1.0, false}};
      (void)blocks;
      // place the border blocks at rows/cols offset by building a 3x3 grid over (n_s+n_u, nh)
      Csr border_col, border_row;
      // pad the u-only border to the full KKT row range [0, n)
      std::vector<int64_t> rp3(n + 1);
      std::vector<int32_t> c3;
      std::vector<double> v3;
      for (int64_t i = 0; i < n; ++i) {
        rp3[i] = (int64_t)c3.size();
        if (i >= P.ns)
          for (int64_t h = 0; h < nh; ++h) {
            c3.push_back((int32_t)h);
            v3.push_back(bh[h * n + i]);
          }
      }
      rp3[n] = (int64_t)c3.size();
      std::vector<int64_t> rp4(nh + 1);
      std::vector<int32_t> c4;
      std::vector<double> v4;
      for (int64_t h = 0; h < nh; ++h) {
        rp4[h] = (int64_t)c4.size();
        for (int64_t i = P.ns; i < n; ++i) {
          c4.push_back((int32_t)i);
          v4.push_back(bh[h * n + i]);
        }
      }
      rp4[nh] = (int64_t)c4.size();
      up(border_col, n, nh, rp3, c3, v3);
      up(border_row, nh, n, rp4, c4, v4);
      BlockRef grid[4] = {{&P.sys.A, 1.0, false}, {&border_col, 1.0, false},
                          {&border_row, 1.0, false}, {nullptr, 1.0, false}};
      int64_t rs[2] = {n, nh};
      csr_block(ctx, 2, grid, rs, rs, *a);
    }
    sizes[0] = P.ns;
    sizes[1] = P.nu;
    sizes[2] = P.nh;
    *out = a;
  });
}

feec_status feec_problem_create(feec_geom* g, feec_bc bc, int grade, const feec_solve_opts* opts,
                                feec_problem** out) {
  return guard([&] {
    FEEC_REQUIRE(g && out, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(abi_bc_ok(bc), FEEC_E_ARG, "bc must be a feec_bc value");
    FEEC_HIP(hipSetDevice(g->mesh->ctx->device));
    auto* P = new feec_problem_s();
    P->g = g;
    P->bc = bc;
    P->k = grade;
    if (opts)
      P->opts = *opts;
    else
      feec_solve_opts_default(&P->opts);
    try {
      validate_opts(P->opts);
      problem_build(*P);
      problem_assemble(*P);
    } catch (...) {
      delete P;
      throw;
    }
    *out = P;
  });
}

void feec_problem_destroy(feec_problem* p) { delete p; }

feec_status feec_problem_assemble(feec_problem* p) {
  return guard([&] { problem_assemble(*p); });
}

feec_status feec_problem_source_buffer(feec_problem* p, double** ptr) {
  return guard([&] { *ptr = p->source.p; });
}

feec_status feec_problem_upload_source(feec_problem* p, const double* src) {
  return guard([&] {
    Mesh& m = *p->g->mesh;
    m.ctx->sync();
    copy_to_device(m.ctx, p->source.p, src, m.n[p->k] * 8);
  });
}

feec_status feec_problem_sizes(const feec_problem* p, int64_t sizes[3], int64_t* nnz) {
  return guard([&] {
    sizes[0] = p->ns;
    sizes[1] = p->nu;
    sizes[2] = p->nh;
    if (nnz) *nnz = p->sys.A.nnz;
  });
}

feec_status feec_problem_preconditioner(const feec_problem* p, int64_t* nnz, int32_t* cheb_degree,
                                        double lambda_max[2], int32_t* column_bytes,
                                        int32_t* format, int64_t* stored, int64_t* offdiag_nnz) {
  return guard([&] {
    FEEC_REQUIRE(p, FEEC_E_ARG, "null problem");
    const KktSystem& s = p->sys;
    const int fmt = kkt_preconditioner_format(s, p->opts.precision, p->opts.precond_storage);
    if (nnz) *nnz = s.P.nnz;
    if (column_bytes) *column_bytes = s.Pcol16.p ? 2 : 4;
    if (format) *format = fmt;
    if (stored)
      *stored = (fmt == FEEC_PC_SELL64_F16 || fmt == FEEC_PC_SELL64_F64) ? s.sell_entries : s.P.nnz;
    if (offdiag_nnz) *offdiag_nnz = s.P_offdiag;
    if (cheb_degree) *cheb_degree = s.cheb_m;
    if (lambda_max) {
      lambda_max[0] = p->sys.lambda_max[0];
      lambda_max[1] = p->sys.lambda_max[1];
    }
    if (s.hx) hx_info(*s.hx, cheb_degree, lambda_max, column_bytes);
  });
}

feec_status feec_problem_apply_preconditioner(feec_problem* p, int32_t precision,
                                              int32_t storage, const double* r, double* z) {
  return guard([&] {
    FEEC_REQUIRE(p, FEEC_E_ARG, "null problem");
    FEEC_REQUIRE(precision >= FEEC_PREC_FP64 && precision <= FEEC_PREC_FP64_PC32, FEEC_E_ARG,
                 "precision must be a feec_precision value");
    FEEC_REQUIRE(storage >= FEEC_PC_STORE_SELL && storage <= FEEC_PC_STORE_CSR, FEEC_E_ARG,
                 "storage must be a feec_pc_storage value");
    FEEC_REQUIRE(precision != FEEC_PREC_FP64_PC32 || storage == FEEC_PC_STORE_SELL, FEEC_E_ARG,
                 "FEEC_PREC_FP64_PC32 runs on the SELL storage only");
    Ctx* ctx = p->g->mesh->ctx;
    const int64_t N = p->sys.n + p->nh;
    DBuf<double> rd(ctx, N), zd(ctx, N);
    copy_to_device(ctx, rd.p, r, N * sizeof(double));
    kkt_apply_preconditioner(p->sys, precision, storage, rd.p, zd.p);
    copy_to_host(ctx, z, zd.p, N * sizeof(double));
  });
}

feec_status feec_problem_solve_device(feec_problem* p, const double* src, feec_report* report) {
  return guard([&] {
    Mesh& m = *p->g->mesh;
    Ctx* ctx = m.ctx;
    const int64_t N = p->sys.n + p->nh;
    FEEC_HIP(hipMemsetAsync(p->rhs.p, 0, N * 8, ctx->stream));
    if (p->bc == FEEC_RELATIVE) m.ensure_interior();
    const int32_t* map = (p->bc == FEEC_RELATIVE) ? m.interior[p->k].p : nullptr;
    k_rhs<<<grid_for(m.n[p->k]), TPB, 0, ctx->stream>>>(src, map, m.n[p->k], p->ns, N, p->rhs.p);
    FEEC_HIP(hipGetLastError());
    feec_report r = kkt_solve(p->sys, p->rhs.p, p->x.p, p->opts, &p->ws);
    p->last = r;
    if (report) *report = r;
    FEEC_REQUIRE(r.converged, FEEC_E_NOT_CONVERGED,
                 "Krylov budget spent before the residual reached rtol");
  });
}

feec_status feec_problem_solve(feec_problem* p, const double* src, feec_report* report) {
  return guard([&] {
    Mesh& m = *p->g->mesh;
    copy_to_device(m.ctx, p->source.p, src, m.n[p->k] * 8);
    feec_status s = feec_problem_solve_device(p, p->source.p, report);
    if (s != FEEC_OK) fail(s, g_last_error);
  });
}

namespace feec {
// b = M_k c on the device (full complex, galerkin.rs:324-371 reading of a
// cochain source), from the problem's element blocks.
static void problem_load(Problem& P, const double* cochain_dev) {
  Mesh& m = *P.g->mesh;
  Ctx* ctx = m.ctx;
  const int k = P.k, nf = P.nf;
  const int64_t nc = m.n[m.d], nk = m.n[k];
  if (!P.have_mass_plan) {
    const int64_t total = nc * nf * nf;
    DBuf<int32_t> rows(ctx, std::max<int64_t>(total, 1)), cols(ctx, std::max<int64_t>(total, 1));
    DBuf<uint32_t> srcs(ctx, std::max<int64_t>(total, 1));
    form_contributions(m, FEEC_FULL, k, k, nf, nf, 0, 0, false, (uint32_t)P.off[5], false, rows.p,
                       cols.p, srcs.p);
    coo_plan_build(ctx, P.mass_plan, nk, nk, total, rows.p, cols.p, srcs.p);
    P.load.alloc(ctx, std::max<int64_t>(nk, 1));
    P.have_mass_plan = true;
  }
  if (P.mass_version != P.el_version) {
    assemble_compacted(ctx, P.mass_plan, P.el.p, P.mass_map, P.Mk);
    P.mass_version = P.el_version;
  }
  spmv_rowseq(ctx, P.Mk, cochain_dev, P.load.p);
}
}  // namespace feec

feec_status feec_problem_solve_cochain_device(feec_problem* p, const double* cochain,
                                              feec_report* report) {
  return guard([&] {
    FEEC_REQUIRE(p && cochain, FEEC_E_ARG, "null argument");
    FEEC_HIP(hipSetDevice(p->g->mesh->ctx->device));
    problem_load(*p, cochain);
    feec_status s = feec_problem_solve_device(p, p->load.p, report);
    if (s != FEEC_OK) fail(s, g_last_error);
  });
}

feec_status feec_problem_solve_cochain(feec_problem* p, const double* cochain, feec_report* report) {
  return guard([&] {
    FEEC_REQUIRE(p && cochain, FEEC_E_ARG, "null argument");
    Mesh& m = *p->g->mesh;
    FEEC_HIP(hipSetDevice(m.ctx->device));
    copy_to_device(m.ctx, p->source.p, cochain, m.n[p->k] * 8);
    feec_status s = feec_problem_solve_cochain_device(p, p->source.p, report);
    if (s != FEEC_OK) fail(s, g_last_error);
  });
}

feec_status feec_problem_solution(feec_problem* p, double* sigma, double* u, double* ph) {
  return guard([&] {
    Mesh& m = *p->g->mesh;
    Ctx* ctx = m.ctx;
    const int k = p->k;
    if (p->bc == FEEC_RELATIVE) m.ensure_interior();
    if (sigma && k >= 1) {
      DBuf<double> o(ctx, m.n[k - 1]);
      const int32_t* map = (p->bc == FEEC_RELATIVE) ? m.interior[k - 1].p : nullptr;
      k_extend<<<grid_for(m.n[k - 1]), TPB, 0, ctx->stream>>>(p->x.p, map, m.n[k - 1], 0, o.p);
      ctx->sync();
      copy_to_host(ctx, sigma, o.p, m.n[k - 1] * 8);
    }
    if (u) {
      DBuf<double> o(ctx, m.n[k]);
      const int32_t* map = (p->bc == FEEC_RELATIVE) ? m.interior[k].p : nullptr;
      k_extend<<<grid_for(m.n[k]), TPB, 0, ctx->stream>>>(p->x.p, map, m.n[k], p->ns, o.p);
      ctx->sync();
      copy_to_host(ctx, u, o.p, m.n[k] * 8);
    }
    if (ph && p->nh) copy_to_host(ctx, ph, p->x.p + p->sys.n, p->nh * 8);
  });
}

feec_status feec_problem_kernel_times(feec_problem* p, int reps, double out[6]) {
  return guard([&] {
    Mesh& m = *p->g->mesh;
    Ctx* ctx = m.ctx;
    hipStream_t st = ctx->stream;
    hipEvent_t e0, e1, e2;
    FEEC_HIP(hipEventCreate(&e0));
    FEEC_HIP(hipEventCreate(&e1));
    FEEC_HIP(hipEventCreate(&e2));
    double* outs[7];
    problem_element_outputs(*p, outs);
    kkt_element_blocks(*p->g, p->k, outs, true);
    DBuf<double> v(ctx, std::max<int64_t>(p->kkt_plan.nnz, 1));
    coo_plan_values(ctx, p->kkt_plan, p->el.p, v.p, nullptr);
    FEEC_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < reps; ++i) kkt_element_blocks(*p->g, p->k, outs, false);
    FEEC_HIP(hipEventRecord(e1, st));
    for (int i = 0; i < reps; ++i) coo_plan_values(ctx, p->kkt_plan, p->el.p, v.p, nullptr);
    FEEC_HIP(hipEventRecord(e2, st));
    FEEC_HIP(hipEventSynchronize(e2));
    float a = 0.f, b = 0.f;
    FEEC_HIP(hipEventElapsedTime(&a, e0, e1));
    FEEC_HIP(hipEventElapsedTime(&b, e1, e2));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    out[0] = a * 1e-3 / reps;
    out[1] = b * 1e-3 / reps;
    out[2] = time_minres_spmv(p->sys, true, reps);
    out[3] = time_minres_spmv(p->sys, false, reps);
    out[4] = time_cheb_step(p->sys, true, reps, p->opts.precond_storage);
    out[5] = time_cheb_step(p->sys, false, reps, p->opts.precond_storage);
  });
}

feec_status feec_problem_mg_kernel_times(feec_problem* p, int reps, int use_graph, double out[4]) {
  return guard([&] {
    FEEC_REQUIRE(p && out, FEEC_E_ARG, "null argument");
    FEEC_REQUIRE(p->sys.hx, FEEC_E_UNSUPPORTED, "the problem's preconditioner is not the multigrid one");
    hx_kernel_times(*p->sys.hx, reps, use_graph != 0, out);
  });
}

feec_status feec_dirac_element_time(feec_geom* g, int reps, double* seconds,
                                   int64_t* values_per_cell) {
  return guard([&] {
    FEEC_REQUIRE(g && seconds, FEEC_E_ARG, "null argument");
    Mesh& m = *g->mesh;
    Ctx* ctx = m.ctx;
    FEEC_HIP(hipSetDevice(ctx->device));
    const int64_t per = dirac_element_count(m.d);
    if (values_per_cell) *values_per_cell = per;
    DBuf<double> out(ctx, std::max<int64_t>(per * m.n[m.d], 1));
    dirac_element_matrices(*g, out.p, true);
    hipEvent_t e0, e1;
    FEEC_HIP(hipEventCreate(&e0));
    FEEC_HIP(hipEventCreate(&e1));
    FEEC_HIP(hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < reps; ++i) dirac_element_matrices(*g, out.p, false);
    FEEC_HIP(hipEventRecord(e1, ctx->stream));
    FEEC_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    FEEC_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *seconds = ms * 1e-3 / std::max(reps, 1);
  });
}

feec_status feec_solve_source(feec_geom* g, feec_bc bc, int grade, const double* src,
                              const feec_solve_opts* opts, double* sigma, double* u, double* ph,
                              feec_report* report) {
  feec_problem* p = nullptr;
  feec_status s = feec_problem_create(g, bc, grade, opts, &p);
  if (s != FEEC_OK) return s;
  s = feec_problem_solve(p, src, report);
  if (s == FEEC_OK || s == FEEC_E_NOT_CONVERGED) {
    feec_status s2 = feec_problem_solution(p, sigma, u, ph);
    if (s == FEEC_OK) s = s2;
  }
  std::string err = g_last_error;
  feec_problem_destroy(p);
  g_last_error = err;
  return s;
}

}  // extern "C"
