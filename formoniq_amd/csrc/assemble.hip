// Scatter-assembly on the device: element matrices -> global CSR.
//
// Reference behaviour (luiswirth/formoniq, crates/):
//   assemble_matrix       formoniq/src/galerkin.rs:138-188 (push v if v != 0.0,
//                         triplets in cell order, COO -> CSR sums duplicates)
//   relative assembly     formoniq/src/whitney_complex.rs:620-624 (E^T A E)
//   hdif_gram             formoniq/src/whitney_complex.rs:180-183 (CSR add)
//   HodgeDirac op         formoniq/src/problems/dirac.rs:242-250 (CSR x CSR)
//
// Every sparse build here is one primitive, the CooPlan: (row, col) keys of
// all contributions are radix-sorted once (stable, so contributions to one
// entry stay in input order, which is cell order for assembly), run heads give
// the CSR entries, and `src` keeps, per entry, the indices of its
// contributors. A value pass is then a deterministic segmented sum over `src`
// with no atomics, so the bits do not depend on scheduling, and re-assembling
// new geometry on a fixed mesh is that pass alone. Bit 31 of a src index
// negates the contribution (the sign flip of the sigma rows of the KKT system).
#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <vector>

#include "feec_internal.hpp"

namespace feec {

namespace {

constexpr int TPB = 256;
constexpr uint32_t NEG = kCooNegate;

inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + TPB - 1) / TPB); }

// Sort keys row * ncols + col (sentinel for dropped contributions) and the
// number of valid ones. Grid-stride with a bounded grid and one atomic per
// block: an atomic per wavefront on the one counter serialized (22 ms for the
// 201 M contributions of a 64^3 KKT system).
constexpr unsigned KEYS_MAX_BLOCKS = 2048;
__global__ void __launch_bounds__(TPB) k_keys(const int32_t* rows, const int32_t* cols, int64_t n,
                                              int64_t ncols, uint64_t sentinel, uint64_t* keys,
                                              unsigned long long* nvalid) {
  __shared__ unsigned long long sh[TPB / 64];
  unsigned long long cnt = 0;  // valid keys of the wavefront, on lane 0
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = rows[i], c = cols[i];
    const bool ok = r >= 0 && c >= 0;
    keys[i] = ok ? (uint64_t)r * (uint64_t)ncols + (uint64_t)c : sentinel;
    // lane 0 is active whenever any lane of its wavefront is
    cnt += (unsigned long long)__popcll(__ballot(ok));
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (unsigned w = 0; w < blockDim.x / 64; ++w) t += sh[w];
    if (t) atomicAdd(nvalid, t);
  }
}

__global__ void k_heads(const uint64_t* keys, int64_t n, int32_t* head) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

__global__ void k_entries(const uint64_t* keys, const int32_t* head, const int32_t* scan, int64_t n,
                          int64_t ncols, int64_t* seg, int32_t* col) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n || !head[i]) return;
  int64_t e = scan[i] - 1;
  seg[e] = i;
  col[e] = (int32_t)(keys[i] % (uint64_t)ncols);
}

// row_ptr from the sorted entries: entry e (or the end, e = nnz) starts every
// row after the previous entry's row up to its own, so each of the nrows + 1
// pointers is written once (empty rows included), without counting atomics.
__global__ void k_row_starts(const uint64_t* keys, const int64_t* seg, int64_t nnz, int64_t ncols,
                             int64_t nrows, int64_t* row_ptr) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e > nnz) return;
  const int64_t prev = e == 0 ? -1 : (int64_t)(keys[seg[e - 1]] / (uint64_t)ncols);
  const int64_t cur = e == nnz ? nrows : (int64_t)(keys[seg[e]] / (uint64_t)ncols);
  for (int64_t r = prev + 1; r <= cur; ++r) row_ptr[r] = e;
}

template <class T>
void exclusive_scan_u64_to_i64(Ctx* ctx, const unsigned long long* in, int64_t* out, int64_t n) {
  size_t bytes = 0;
  FEEC_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, (int64_t)0, (size_t)n,
                                   rocprim::plus<int64_t>(), ctx->stream));
  void* tmp = ctx->ensure_temp(bytes);
  FEEC_HIP(rocprim::exclusive_scan(tmp, bytes, in, out, (int64_t)0, (size_t)n,
                                   rocprim::plus<int64_t>(), ctx->stream));
}

void inclusive_scan_i32(Ctx* ctx, const int32_t* in, int32_t* out, int64_t n) {
  size_t bytes = 0;
  FEEC_HIP(rocprim::inclusive_scan(nullptr, bytes, in, out, (size_t)n, rocprim::plus<int32_t>(),
                                   ctx->stream));
  void* tmp = ctx->ensure_temp(bytes);
  FEEC_HIP(rocprim::inclusive_scan(tmp, bytes, in, out, (size_t)n, rocprim::plus<int32_t>(),
                                   ctx->stream));
}

// One thread per CSR entry: sum contributors in order; keep = some contributor != 0.
__global__ void k_segsum(const int64_t* seg, const uint32_t* src, int64_t nnz, const double* vals,
                         double* out, uint8_t* keep, unsigned long long* ndropped) {
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  int64_t b = seg[e], en = seg[e + 1];
  double acc = 0.0;
  bool nz = false;
  for (int64_t s = b; s < en; ++s) {
    uint32_t j = src[s];
    double v = vals[j & ~NEG];
    if (j & NEG) v = -v;
    nz |= (v != 0.0);
    acc = (s == b) ? v : acc + v;
  }
  out[e] = acc;
  if (keep) {
    keep[e] = nz ? 1 : 0;
    if (!nz) atomicAdd(ndropped, 1ull);
  }
}

// The same sums with the gathers shared by the wavefront: each wavefront
// takes 64 consecutive entries, whose contributions are contiguous in src;
// its lanes fetch all of them together (every lane busy, 64 gathers in
// flight) into LDS, then each lane adds its own entry's values in order, so
// every sum is the one k_segsum forms. A run longer than the LDS stage falls
// back to the per-lane gathers.
// XCD_CHUNK > 0: blocks are dealt to the 8 XCDs round-robin; the grid (a
// multiple of 8 * XCD_CHUNK) is renumbered so that each XCD takes runs of
// XCD_CHUNK consecutive blocks spread over the whole range, so that an
// element value's nearby readers share that XCD's L2.
// STAGE: contributions staged per wavefront (a smaller stage leaves room for
// more resident wavefronts; runs longer than it take the per-lane gathers).
template <int XCD_CHUNK, int STAGE>
__global__ void __launch_bounds__(TPB) k_segsum_wave(const int64_t* seg, const uint32_t* src, int64_t nnz,
                                                     const double* vals, double* out, uint8_t* keep,
                                                     unsigned long long* ndropped) {
  __shared__ double sv[TPB / 64][STAGE];
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  int64_t blk = blockIdx.x;
  if (XCD_CHUNK > 0) {
    const int64_t x = blockIdx.x % 8, k = blockIdx.x / 8;
    blk = ((k / XCD_CHUNK) * 8 + x) * XCD_CHUNK + k % XCD_CHUNK;
  }
  const int64_t e0 = (blk * (TPB / 64) + w) * 64;
  const bool active = e0 < nnz;
  const int64_t e1 = active ? (e0 + 64 < nnz ? e0 + 64 : nnz) : e0;
  const int64_t s0 = active ? seg[e0] : 0, s1 = active ? seg[e1] : 0;
  const bool staged = active && s1 - s0 <= STAGE;
  if (staged)
    for (int64_t q = s0 + lane; q < s1; q += 64) {
      const uint32_t j = src[q];
      const double v = vals[j & ~NEG];
      sv[w][q - s0] = (j & NEG) ? -v : v;
    }
  __syncthreads();
  const int64_t e = e0 + lane;
  if (!active || e >= e1) return;
  const int64_t b = seg[e], en = seg[e + 1];
  double acc = 0.0;
  bool nz = false;
  for (int64_t q = b; q < en; ++q) {
    double v;
    if (staged) {
      v = sv[w][q - s0];
    } else {
      const uint32_t j = src[q];
      v = vals[j & ~NEG];
      if (j & NEG) v = -v;
    }
    nz |= (v != 0.0);
    acc = (q == b) ? v : acc + v;
  }
  out[e] = acc;
  if (keep) {
    keep[e] = nz ? 1 : 0;
    if (!nz) atomicAdd(ndropped, 1ull);
  }
}

__global__ void k_rows_of(const int64_t* row_ptr, int64_t nrows, int32_t* rows) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  for (int64_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) rows[p] = (int32_t)r;
}

// Compaction of a pattern to its kept entries. Entry p moves to newpos[p]
// (the exclusive scan of keep), one lane per entry so loads and stores run in
// entry order; a row starts at the new position of its first old entry, or at
// the end when it has none left in the old array.
__global__ void k_compact_entries(const int32_t* col, const double* val, const uint8_t* keep,
                                  const int64_t* newpos, int64_t n, int32_t* ocol, double* oval) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p < n && keep[p]) {
    ocol[newpos[p]] = col[p];
    oval[newpos[p]] = val[p];
  }
}
__global__ void k_compact_rows(const int64_t* row_ptr, int64_t nrows, const int64_t* newpos,
                               int64_t n, int64_t nnz_kept, int64_t* orow) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r > nrows) return;
  const int64_t p = row_ptr[r];
  orow[r] = p < n ? newpos[p] : nnz_kept;
}

__global__ void k_u8_to_u64(const uint8_t* a, int64_t n, unsigned long long* b) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_set_last(int64_t* a, int64_t idx, int64_t v) { a[idx] = v; }

template <class T>
__global__ void k_copy_scale(const T* a, int64_t n, double s, T* b) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_diag(const int64_t* row_ptr, const int32_t* col, const double* val, int64_t n,
                       double* diag) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  double d = 0.0;
  for (int64_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p)
    if (col[p] == r) d = val[p];
  diag[r] = d;
}

}  // namespace

void coo_plan_build(Ctx* ctx, CooPlan& plan, int64_t nrows, int64_t ncols, int64_t n,
                    const int32_t* rows, const int32_t* cols, const uint32_t* srcs) {
  plan.nrows = nrows;
  plan.ncols = ncols;
  FEEC_REQUIRE(nrows >= 0 && ncols >= 0, FEEC_E_ARG, "negative matrix shape");
  const uint64_t sentinel = (uint64_t)nrows * (uint64_t)std::max<int64_t>(ncols, 1);
  int end_bit = 1;
  while (end_bit < 64 && (uint64_t(1) << end_bit) <= sentinel) ++end_bit;
  plan.row_ptr.alloc(ctx, nrows + 1);
  if (n == 0 || nrows == 0 || ncols == 0) {
    FEEC_HIP(hipMemsetAsync(plan.row_ptr.p, 0, (nrows + 1) * sizeof(int64_t), ctx->stream));
    plan.nnz = 0;
    plan.col.alloc(ctx, 0);
    plan.seg.alloc(ctx, 1);
    FEEC_HIP(hipMemsetAsync(plan.seg.p, 0, sizeof(int64_t), ctx->stream));
    plan.src.alloc(ctx, 0);
    ctx->sync();
    return;
  }
  DBuf<uint64_t> k0(ctx, n), k1(ctx, n);
  DBuf<unsigned long long> nvalid(ctx, 1);
  FEEC_HIP(hipMemsetAsync(nvalid.p, 0, sizeof(unsigned long long), ctx->stream));
  k_keys<<<std::min(grid_for(n), KEYS_MAX_BLOCKS), TPB, 0, ctx->stream>>>(
      rows, cols, n, std::max<int64_t>(ncols, 1), sentinel, k0.p, nvalid.p);
  DBuf<uint32_t> s0(ctx, n);
  FEEC_HIP(hipMemcpyAsync(s0.p, srcs, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
  plan.src.alloc(ctx, n);
  {
    size_t bytes = 0;
    FEEC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, k0.p, k1.p, s0.p, plan.src.p, (size_t)n, 0,
                                       (unsigned)end_bit, ctx->stream));
    void* tmp = ctx->ensure_temp(bytes);
    FEEC_HIP(rocprim::radix_sort_pairs(tmp, bytes, k0.p, k1.p, s0.p, plan.src.p, (size_t)n, 0,
                                       (unsigned)end_bit, ctx->stream));
  }
  k0.reset();
  s0.reset();
  unsigned long long hv = 0;
  FEEC_HIP(hipMemcpyAsync(&hv, nvalid.p, sizeof(hv), hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  int64_t nv = (int64_t)hv;
  FEEC_REQUIRE(nv < (int64_t(1) << 31), FEEC_E_ARG, "too many contributions for one matrix");
  DBuf<int32_t> head(ctx, std::max<int64_t>(nv, 1)), scan(ctx, std::max<int64_t>(nv, 1));
  int64_t nnz = 0;
  if (nv > 0) {
    k_heads<<<grid_for(nv), TPB, 0, ctx->stream>>>(k1.p, nv, head.p);
    inclusive_scan_i32(ctx, head.p, scan.p, nv);
    int32_t last = 0;
    FEEC_HIP(hipMemcpyAsync(&last, scan.p + nv - 1, sizeof(int32_t), hipMemcpyDeviceToHost,
                            ctx->stream));
    ctx->sync();
    nnz = last;
  }
  plan.nnz = nnz;
  plan.seg.alloc(ctx, nnz + 1);
  plan.col.alloc(ctx, nnz);
  if (nv > 0)
    k_entries<<<grid_for(nv), TPB, 0, ctx->stream>>>(k1.p, head.p, scan.p, nv,
                                                     std::max<int64_t>(ncols, 1), plan.seg.p,
                                                     plan.col.p);
  k_set_last<<<1, 1, 0, ctx->stream>>>(plan.seg.p, nnz, nv);
  k_row_starts<<<grid_for(nnz + 1), TPB, 0, ctx->stream>>>(k1.p, plan.seg.p, nnz, std::max<int64_t>(ncols, 1),
                                                           nrows, plan.row_ptr.p);
  FEEC_HIP(hipGetLastError());
  ctx->sync();
}

int64_t coo_plan_values(Ctx* ctx, const CooPlan& plan, const double* vals, double* out,
                        uint8_t* keep) {
  if (plan.nnz == 0) return 0;
  unsigned long long* nd = ctx->dcount;
  // per-lane gathers, or the wavefront's shared ones in XCD runs with a stage
  // of 512 or 1,024 contributions: the same sums, bit for bit. Below 2^22 entries the per-lane gathers were the fastest at every
  // size measured, so only larger plans are timed.
  const bool timed = plan.nnz >= (int64_t(1) << 22);
  tune_run(ctx, "segsum", 0, 0, plan.nnz, 3, [&](int c) {
    if (keep) FEEC_HIP(hipMemsetAsync(nd, 0, sizeof(unsigned long long), ctx->stream));
    const unsigned g = grid_for(plan.nnz);
    constexpr int C = 128;
    const unsigned gx = (g + 8 * C - 1) / (8 * C) * (8 * C);
    if (c == 0)
      k_segsum<<<g, TPB, 0, ctx->stream>>>(plan.seg.p, plan.src.p, plan.nnz, vals, out, keep, nd);
    else if (c == 1)  // about 2-3 contributions per entry (the KKT)
      k_segsum_wave<C, 512><<<gx, TPB, 0, ctx->stream>>>(plan.seg.p, plan.src.p, plan.nnz, vals, out, keep, nd);
    else  // about 5 (mass plus stiffness: the preconditioner matrix)
      k_segsum_wave<C, 1024><<<gx, TPB, 0, ctx->stream>>>(plan.seg.p, plan.src.p, plan.nnz, vals, out, keep, nd);
    FEEC_HIP(hipGetLastError());
  }, timed);
  if (!keep) return 0;
  unsigned long long h = 0;
  FEEC_HIP(hipMemcpyAsync(&h, nd, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  return (int64_t)h;
}

void coo_plan_apply(Ctx* ctx, const CooPlan& plan, const double* vals, bool zero_drop, Csr& out) {
  out.ctx = ctx;
  out.nrows = plan.nrows;
  out.ncols = plan.ncols;
  DBuf<double> v(ctx, plan.nnz);
  DBuf<uint8_t> keep;
  if (zero_drop) keep.alloc(ctx, plan.nnz);
  int64_t dropped = coo_plan_values(ctx, plan, vals, v.p, zero_drop ? keep.p : nullptr);
  if (dropped == 0) {
    out.nnz = plan.nnz;
    out.row_ptr.alloc(ctx, plan.nrows + 1);
    out.col.alloc(ctx, plan.nnz);
    FEEC_HIP(hipMemcpyAsync(out.row_ptr.p, plan.row_ptr.p, (plan.nrows + 1) * sizeof(int64_t),
                            hipMemcpyDeviceToDevice, ctx->stream));
    FEEC_HIP(hipMemcpyAsync(out.col.p, plan.col.p, plan.nnz * sizeof(int32_t),
                            hipMemcpyDeviceToDevice, ctx->stream));
    out.val = std::move(v);
    ctx->sync();
    return;
  }
  // Entries whose contributions were all exactly zero were never pushed by
  // the reference: compact them away.
  int64_t nnz = plan.nnz - dropped;
  DBuf<unsigned long long> k64(ctx, plan.nnz);
  k_u8_to_u64<<<grid_for(plan.nnz), TPB, 0, ctx->stream>>>(keep.p, plan.nnz, k64.p);
  DBuf<int64_t> pos(ctx, plan.nnz);
  exclusive_scan_u64_to_i64<int64_t>(ctx, k64.p, pos.p, plan.nnz);
  out.nnz = nnz;
  out.col.alloc(ctx, nnz);
  out.val.alloc(ctx, nnz);
  out.row_ptr.alloc(ctx, plan.nrows + 1);
  k_compact_entries<<<grid_for(plan.nnz), TPB, 0, ctx->stream>>>(plan.col.p, v.p, keep.p, pos.p,
                                                                 plan.nnz, out.col.p, out.val.p);
  k_compact_rows<<<grid_for(plan.nrows + 1), TPB, 0, ctx->stream>>>(
      plan.row_ptr.p, plan.nrows, pos.p, plan.nnz, nnz, out.row_ptr.p);
  FEEC_HIP(hipGetLastError());
  ctx->sync();
}

// ---------------------------------------------------------------------------
// sparse utilities
// ---------------------------------------------------------------------------
namespace {

__global__ void k_iota_u32(uint32_t* a, int64_t n, uint32_t base) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = base + (uint32_t)i;
}

__global__ void k_block_coo(const int64_t* row_ptr, const int32_t* col, int64_t nrows,
                            int64_t roff, int64_t coff, bool transpose, uint32_t src_base,
                            bool negate, int32_t* orow, int32_t* ocol, uint32_t* osrc) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  for (int64_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
    int32_t a = (int32_t)r, b = col[p];
    orow[p] = (int32_t)((transpose ? b : a) + roff);
    ocol[p] = (int32_t)((transpose ? a : b) + coff);
    osrc[p] = (src_base + (uint32_t)p) | (negate ? NEG : 0u);
  }
}

__global__ void k_scale_copy(const double* a, int64_t n, double s, double* b) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (s == 1.0 || s == -1.0) ? a[i] : s * a[i];
}

__global__ void k_select_coo(const int64_t* row_ptr, const int32_t* col, int64_t nrows,
                             const int32_t* rmap, const int32_t* cmap, int32_t* orow,
                             int32_t* ocol, uint32_t* osrc) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  int32_t nr = rmap ? rmap[r] : (int32_t)r;
  for (int64_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
    orow[p] = nr;
    ocol[p] = cmap ? cmap[col[p]] : col[p];
    osrc[p] = (uint32_t)p;
  }
}

// structural product: count terms per row of A
__global__ void k_mul_count(const int64_t* arp, const int32_t* acol, const int64_t* brp,
                            int64_t nrows, unsigned long long* cnt) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  unsigned long long c = 0;
  for (int64_t p = arp[r]; p < arp[r + 1]; ++p) c += brp[acol[p] + 1] - brp[acol[p]];
  cnt[r] = c;
}

__global__ void k_mul_terms(const int64_t* arp, const int32_t* acol, const double* aval,
                            const int64_t* brp, const int32_t* bcol, const double* bval,
                            int64_t nrows, const int64_t* start, int32_t* orow, int32_t* ocol,
                            uint32_t* osrc, double* oval) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  int64_t t = start[r];
  for (int64_t p = arp[r]; p < arp[r + 1]; ++p) {
    int32_t k = acol[p];
    double av = aval[p];
    for (int64_t q = brp[k]; q < brp[k + 1]; ++q) {
      orow[t] = (int32_t)r;
      ocol[t] = bcol[q];
      osrc[t] = (uint32_t)t;
      oval[t] = av * bval[q];
      ++t;
    }
  }
}

}  // namespace

void csr_block(Ctx* ctx, int nb, const BlockRef* blocks, const int64_t* row_sizes,
               const int64_t* col_sizes, Csr& out) {
  std::vector<int64_t> roff(nb + 1, 0), coff(nb + 1, 0);
  for (int i = 0; i < nb; ++i) {
    roff[i + 1] = roff[i] + row_sizes[i];
    coff[i + 1] = coff[i] + col_sizes[i];
  }
  int64_t total = 0;
  for (int i = 0; i < nb * nb; ++i)
    if (blocks[i].a) total += blocks[i].a->nnz;
  DBuf<int32_t> rows(ctx, total), cols(ctx, total);
  DBuf<uint32_t> srcs(ctx, total);
  DBuf<double> vals(ctx, total);
  int64_t at = 0;
  for (int bi = 0; bi < nb; ++bi)
    for (int bj = 0; bj < nb; ++bj) {
      const BlockRef& b = blocks[bi * nb + bj];
      if (!b.a || b.a->nnz == 0) continue;
      double s = b.scale;
      k_block_coo<<<grid_for(b.a->nrows), TPB, 0, ctx->stream>>>(
          b.a->row_ptr.p, b.a->col.p, b.a->nrows, roff[bi], coff[bj], b.transpose,
          (uint32_t)at, s == -1.0, rows.p + at, cols.p + at, srcs.p + at);
      k_scale_copy<<<grid_for(b.a->nnz), TPB, 0, ctx->stream>>>(b.a->val.p, b.a->nnz, s,
                                                                vals.p + at);
      at += b.a->nnz;
    }
  CooPlan plan;
  coo_plan_build(ctx, plan, roff[nb], coff[nb], total, rows.p, cols.p, srcs.p);
  coo_plan_apply(ctx, plan, vals.p, false, out);
}

// Exclusive prefix sum of n counts (n >= 1) into out[0..n).
void exclusive_scan_counts(Ctx* ctx, const unsigned long long* in, int64_t* out, int64_t n) {
  exclusive_scan_u64_to_i64<int64_t>(ctx, in, out, n);
}

void csr_transpose(const Csr& a, Csr& out) {
  BlockRef b{&a, 1.0, true};
  int64_t r = a.ncols, c = a.nrows;
  csr_block(a.ctx, 1, &b, &r, &c, out);
}

void csr_add(const Csr& a, const Csr& b, Csr& out) {
  Ctx* ctx = a.ctx;
  int64_t total = a.nnz + b.nnz;
  DBuf<int32_t> rows(ctx, total), cols(ctx, total);
  DBuf<uint32_t> srcs(ctx, total);
  DBuf<double> vals(ctx, total);
  if (a.nnz) {
    k_block_coo<<<grid_for(a.nrows), TPB, 0, ctx->stream>>>(a.row_ptr.p, a.col.p, a.nrows, 0, 0,
                                                            false, 0, false, rows.p, cols.p, srcs.p);
    FEEC_HIP(hipMemcpyAsync(vals.p, a.val.p, a.nnz * sizeof(double), hipMemcpyDeviceToDevice,
                            ctx->stream));
  }
  if (b.nnz) {
    k_block_coo<<<grid_for(b.nrows), TPB, 0, ctx->stream>>>(
        b.row_ptr.p, b.col.p, b.nrows, 0, 0, false, (uint32_t)a.nnz, false, rows.p + a.nnz,
        cols.p + a.nnz, srcs.p + a.nnz);
    FEEC_HIP(hipMemcpyAsync(vals.p + a.nnz, b.val.p, b.nnz * sizeof(double),
                            hipMemcpyDeviceToDevice, ctx->stream));
  }
  CooPlan plan;
  coo_plan_build(ctx, plan, a.nrows, a.ncols, total, rows.p, cols.p, srcs.p);
  coo_plan_apply(ctx, plan, vals.p, false, out);
}

void csr_select(const Csr& a, const int32_t* row_map, int64_t nrows, const int32_t* col_map,
                int64_t ncols, Csr& out) {
  Ctx* ctx = a.ctx;
  DBuf<int32_t> rows(ctx, a.nnz), cols(ctx, a.nnz);
  DBuf<uint32_t> srcs(ctx, a.nnz);
  if (a.nnz)
    k_select_coo<<<grid_for(a.nrows), TPB, 0, ctx->stream>>>(a.row_ptr.p, a.col.p, a.nrows,
                                                             row_map, col_map, rows.p, cols.p,
                                                             srcs.p);
  CooPlan plan;
  coo_plan_build(ctx, plan, nrows, ncols, a.nnz, rows.p, cols.p, srcs.p);
  coo_plan_apply(ctx, plan, a.val.p, false, out);
}

void csr_mul(const Csr& a, const Csr& b, Csr& out) {
  Ctx* ctx = a.ctx;
  FEEC_REQUIRE(a.ncols == b.nrows, FEEC_E_ARG, "csr_mul shape mismatch");
  DBuf<unsigned long long> cnt(ctx, std::max<int64_t>(a.nrows, 1));
  DBuf<int64_t> start(ctx, std::max<int64_t>(a.nrows, 1));
  if (a.nrows)
    k_mul_count<<<grid_for(a.nrows), TPB, 0, ctx->stream>>>(a.row_ptr.p, a.col.p, b.row_ptr.p,
                                                            a.nrows, cnt.p);
  int64_t total = 0;
  if (a.nrows) {
    exclusive_scan_u64_to_i64<int64_t>(ctx, cnt.p, start.p, a.nrows);
    int64_t last_start = 0;
    unsigned long long last_cnt = 0;
    FEEC_HIP(hipMemcpyAsync(&last_start, start.p + a.nrows - 1, sizeof(int64_t),
                            hipMemcpyDeviceToHost, ctx->stream));
    FEEC_HIP(hipMemcpyAsync(&last_cnt, cnt.p + a.nrows - 1, sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    total = last_start + (int64_t)last_cnt;
  }
  FEEC_REQUIRE(total < (int64_t(1) << 31), FEEC_E_ARG, "product too large");
  DBuf<int32_t> rows(ctx, total), cols(ctx, total);
  DBuf<uint32_t> srcs(ctx, total);
  DBuf<double> vals(ctx, total);
  if (a.nrows)
    k_mul_terms<<<grid_for(a.nrows), TPB, 0, ctx->stream>>>(
        a.row_ptr.p, a.col.p, a.val.p, b.row_ptr.p, b.col.p, b.val.p, a.nrows, start.p, rows.p,
        cols.p, srcs.p, vals.p);
  CooPlan plan;
  coo_plan_build(ctx, plan, a.nrows, b.ncols, total, rows.p, cols.p, srcs.p);
  coo_plan_apply(ctx, plan, vals.p, false, out);
}

void csr_diagonal(const Csr& a, double* diag_dev) {
  int64_t n = std::min(a.nrows, a.ncols);
  k_diag<<<grid_for(n), TPB, 0, a.ctx->stream>>>(a.row_ptr.p, a.col.p, a.val.p, n, diag_dev);
  FEEC_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// element matrices -> Galerkin matrix
// ---------------------------------------------------------------------------
namespace {

// contributions of one element block, cell-major (t = cell * r * c + i * c + j),
// source index = SoA slot (i * c + j) * ncells + cell + base
__global__ void k_form_coo(const int32_t* faces_r, int nfr, const int32_t* faces_c, int nfc,
                           const int32_t* rmap, const int32_t* cmap, int64_t roff, int64_t coff,
                           int64_t nc, int r, int c, bool transpose, uint32_t base, bool negate,
                           int32_t* orow, int32_t* ocol, uint32_t* osrc) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t per = (int64_t)r * c;
  if (t >= nc * per) return;
  int64_t cell = t / per;
  int e = (int)(t % per);
  int i = e / c, j = e % c;
  int32_t gi = faces_r[cell * nfr + i];
  int32_t gj = faces_c[cell * nfc + j];
  if (rmap) gi = rmap[gi];
  if (cmap) gj = cmap[gj];
  bool ok = gi >= 0 && gj >= 0;
  int32_t R = ok ? (int32_t)(gi + roff) : -1, C = ok ? (int32_t)(gj + coff) : -1;
  orow[t] = transpose ? C : R;
  ocol[t] = transpose ? R : C;
  osrc[t] = (base + (uint32_t)((int64_t)e * nc + cell)) | (negate ? NEG : 0u);
}

}  // namespace

struct FormSide {
  int grade;
  int64_t offset;
};

// Contribution list of one element block of size r x c whose rows are the
// local faces of grade rg and columns of grade cg.
void form_contributions(Mesh& m, int bc, int rg, int cg, int r, int c, int64_t roff, int64_t coff,
                        bool transpose, uint32_t base, bool negate, int32_t* rows, int32_t* cols,
                        uint32_t* srcs) {
  int64_t nc = m.n[m.d];
  if (bc == FEEC_RELATIVE) m.ensure_interior();
  const int32_t* rmap = bc == FEEC_RELATIVE ? m.interior[rg].p : nullptr;
  const int32_t* cmap = bc == FEEC_RELATIVE ? m.interior[cg].p : nullptr;
  k_form_coo<<<grid_for(nc * r * c), TPB, 0, m.ctx->stream>>>(
      m.faces[rg].p, m.nfaces(rg), m.faces[cg].p, m.nfaces(cg), rmap, cmap, roff, coff, nc, r, c,
      transpose, base, negate, rows, cols, srcs);
  FEEC_HIP(hipGetLastError());
}

int64_t ndofs(Mesh& m, int bc, int k) {
  if (k < 0 || k > m.d) return 0;
  if (bc == FEEC_RELATIVE) {
    m.ensure_interior();
    return m.n_interior[k];
  }
  return m.n[k];
}

const CooPlan& form_plan(Mesh& m, int bc, int kind, int grade) {
  auto key = std::make_tuple(bc, kind, grade);
  {
    std::lock_guard<std::mutex> lock(m.mu);
    auto it = m.plans.find(key);
    if (it != m.plans.end()) return *it->second;
  }
  int tg = (kind == 2 || kind == 3) ? grade - 1 : grade;
  int rg = (kind == 1 || kind == 3) ? grade - 1 : grade;
  int r = element_rows(m.d, kind, grade), c = element_cols(m.d, kind, grade);
  int64_t nc = m.n[m.d];
  int64_t n = nc * r * c;
  auto plan = std::make_unique<CooPlan>();
  {
    DBuf<int32_t> rows(m.ctx, n), cols(m.ctx, n);
    DBuf<uint32_t> srcs(m.ctx, n);
    form_contributions(m, bc, tg, rg, r, c, 0, 0, false, 0, false, rows.p, cols.p, srcs.p);
    coo_plan_build(m.ctx, *plan, ndofs(m, bc, tg), ndofs(m, bc, rg), n, rows.p, cols.p, srcs.p);
  }
  std::lock_guard<std::mutex> lock(m.mu);
  auto& slot = m.plans[key];
  slot = std::move(plan);
  return *slot;
}

void assemble_form(Geom& g, int bc, int kind, int grade, Csr& out) {
  Mesh& m = *g.mesh;
  Ctx* ctx = m.ctx;
  out.ctx = ctx;
  if (kind == FEEC_HDIF_GRAM) {
    Csr a, b;
    assemble_form(g, bc, FEEC_MASS, grade, a);
    assemble_form(g, bc, FEEC_DIF_BOTH, grade + 1, b);
    csr_add(a, b, out);
    return;
  }
  int tg = (kind == 2 || kind == 3) ? grade - 1 : grade;
  int rg = (kind == 1 || kind == 3) ? grade - 1 : grade;
  int64_t nr = ndofs(m, bc, tg), ncl = ndofs(m, bc, rg);
  int r = element_rows(m.d, kind, grade), c = element_cols(m.d, kind, grade);
  // HilbertComplex::pairing (whitney_complex.rs:113-122): empty off the complex
  if (grade < 0 || grade > m.d || nr == 0 || ncl == 0 || r == 0 || c == 0) {
    out.nrows = nr;
    out.ncols = ncl;
    out.nnz = 0;
    out.row_ptr.alloc(ctx, nr + 1);
    FEEC_HIP(hipMemsetAsync(out.row_ptr.p, 0, (nr + 1) * sizeof(int64_t), ctx->stream));
    out.col.alloc(ctx, 0);
    out.val.alloc(ctx, 0);
    ctx->sync();
    return;
  }
  int64_t nc = m.n[m.d];
  DBuf<double> el(ctx, (size_t)nc * r * c);
  element_matrices_soa(g, kind, grade, el.p);
  const CooPlan& plan = form_plan(m, bc, kind, grade);
  coo_plan_apply(ctx, plan, el.p, true, out);
}

}  // namespace feec

namespace feec {

namespace {
__global__ void k_scatter_kept(const double* vals, const uint8_t* keep, const int64_t* pos,
                               int64_t n, double* out) {
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < n && keep[e]) out[pos[e]] = vals[e];
}
__global__ void k_keep_diff(const uint8_t* a, const uint8_t* b, int64_t n,
                            unsigned long long* cnt) {
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < n && a[e] != b[e]) atomicAdd(cnt, 1ull);
}
__global__ void k_dense_column(const int64_t* row_ptr, int64_t nrows, const double* v,
                               double* out) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  int64_t b = row_ptr[r];
  out[r] = (row_ptr[r + 1] > b) ? v[b] : 0.0;
}
}  // namespace

void coo_plan_dense_column(Ctx* ctx, const CooPlan& plan, const double* vals, double* out) {
  DBuf<double> v(ctx, std::max<int64_t>(plan.nnz, 1));
  coo_plan_values(ctx, plan, vals, v.p, nullptr);
  k_dense_column<<<grid_for(plan.nrows), TPB, 0, ctx->stream>>>(plan.row_ptr.p, plan.nrows, v.p,
                                                                out);
  FEEC_HIP(hipGetLastError());
}

// Values of a plan into a CSR whose pattern is the plan's minus the entries
// whose contributions are all exactly zero. The first call (or a call where the
// set of such entries changed) rebuilds the compacted pattern and returns true.
bool assemble_compacted(Ctx* ctx, const CooPlan& plan, const double* vals, CompactMap& cm,
                        Csr& A) {
  const int64_t ns = plan.nnz;
  DBuf<double> v(ctx, ns);
  DBuf<uint8_t> keep(ctx, ns);
  int64_t dropped = coo_plan_values(ctx, plan, vals, v.p, keep.p);
  bool rebuild = !cm.valid;
  if (cm.valid) {
    DBuf<unsigned long long> cnt(ctx, 1);
    FEEC_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), ctx->stream));
    if (ns)
      k_keep_diff<<<grid_for(ns), TPB, 0, ctx->stream>>>(keep.p, cm.keep.p, ns, cnt.p);
    unsigned long long h = 0;
    FEEC_HIP(hipMemcpyAsync(&h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    rebuild = h != 0;
  }
  if (rebuild) {
    A.ctx = ctx;
    A.nrows = plan.nrows;
    A.ncols = plan.ncols;
    A.nnz = ns - dropped;
    A.row_ptr.alloc(ctx, plan.nrows + 1);
    A.col.alloc(ctx, A.nnz);
    A.val.alloc(ctx, A.nnz);
    cm.pos.alloc(ctx, ns);
    if (ns) {
      DBuf<unsigned long long> k64(ctx, ns);
      k_u8_to_u64<<<grid_for(ns), TPB, 0, ctx->stream>>>(keep.p, ns, k64.p);
      exclusive_scan_u64_to_i64<int64_t>(ctx, k64.p, cm.pos.p, ns);
    }
    if (ns)
      k_compact_entries<<<grid_for(ns), TPB, 0, ctx->stream>>>(plan.col.p, v.p, keep.p, cm.pos.p,
                                                               ns, A.col.p, A.val.p);
    k_compact_rows<<<grid_for(plan.nrows + 1), TPB, 0, ctx->stream>>>(
        plan.row_ptr.p, plan.nrows, cm.pos.p, ns, A.nnz, A.row_ptr.p);
    cm.keep = std::move(keep);
    cm.valid = true;
  } else if (ns) {
    k_scatter_kept<<<grid_for(ns), TPB, 0, ctx->stream>>>(v.p, cm.keep.p, cm.pos.p, ns, A.val.p);
  }
  FEEC_HIP(hipGetLastError());
  ctx->sync();
  return rebuild;
}

}  // namespace feec
