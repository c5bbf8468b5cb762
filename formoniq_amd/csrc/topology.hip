// Topology on the device: colex numbering of every skeleton by radix sort,
// the cell -> face maps, boundary flags, coboundary operators and coordinates
// -> squared edge lengths.
//
// Reference behaviour (luiswirth/formoniq, crates/):
//   Complex::from_cells_unchecked   simplicial/src/topology/complex.rs:298-336
//   Skeleton::new (sort + dedup)    simplicial/src/topology/skeleton.rs:50-83
//   Simplex colex Ord               simplicial/src/topology/simplex.rs:126-130
//   boundary_simplices              simplicial/src/topology/complex.rs:140-149
//   incidences / coboundary         simplicial/src/topology/complex.rs:168-217
//   to_edge_lengths_sq              regge/src/coord/mesh.rs:208-216
//
// Design: a k-simplex is the sorted vertex tuple (v_0 < ... < v_k); its colex
// order is the lexicographic order of (v_k, ..., v_0). Tuples are packed
// several vertices per 64-bit word, highest position in the highest bits, and
// sorted least-significant word first with a stable LSD radix sort, so ties
// keep input order and the result is deterministic. Numbering is then a scan
// over run heads; every cell writes the global index of each local face
// straight from the sorted permutation, with no hashing.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <numeric>

#include "feec_internal.hpp"

namespace feec {

namespace {

constexpr int TPB = 256;

inline unsigned grid_for(int64_t n, int tpb = TPB) {
  return (unsigned)std::max<int64_t>(1, (n + tpb - 1) / tpb);
}

// Local colex (k+1)-subsets of the d+1 cell vertices (simplex.rs:184-193).
struct LocalSubsets {
  int count = 0;
  int size = 0;
  uint8_t elem[10][5];  // [subset][position]
};

LocalSubsets local_subsets(int d, int k) {
  LocalSubsets t;
  t.size = k + 1;
  for (unsigned m = 0; m < (1u << (d + 1)); ++m) {
    if (__builtin_popcount(m) != k + 1) continue;
    int p = 0;
    for (int e = 0; e <= d; ++e)
      if (m & (1u << e)) t.elem[t.count][p++] = (uint8_t)e;
    ++t.count;
  }
  return t;
}

struct Packing {
  int bits = 1;      // bits per vertex id
  int per_word = 1;  // vertices per 64-bit word
  int words = 1;
};

Packing packing(int64_t nverts, int tuple) {
  Packing p;
  p.bits = 1;
  while ((int64_t(1) << p.bits) < nverts) ++p.bits;
  p.per_word = std::max(1, 64 / p.bits);
  p.words = (tuple + p.per_word - 1) / p.per_word;
  return p;
}

// word w of the tuple of slot s = cell * nf + f (cells row-major, d+1 each)
__device__ inline uint64_t tuple_word(const int32_t* cells, int nv, const LocalSubsets& ls,
                                      int64_t slot, int w, int bits, int per_word) {
  int nf = ls.count;
  int64_t cell = slot / nf;
  int f = (int)(slot % nf);
  uint64_t key = 0;
  int lo = w * per_word;
  int hi = min(lo + per_word, ls.size);
  for (int p = hi - 1; p >= lo; --p) key = (key << bits) | (uint64_t)cells[cell * nv + ls.elem[f][p]];
  return key;
}

template <class T>
__global__ void k_iota(T* a, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = (T)i;
}

__global__ void k_gather_words(const int32_t* cells, int nv, LocalSubsets ls, const uint32_t* perm,
                               int64_t n, int w, int bits, int per_word, uint64_t* keys) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) keys[i] = tuple_word(cells, nv, ls, perm[i], w, bits, per_word);
}

__device__ inline bool same_tuple(const int32_t* cells, int nv, const LocalSubsets& ls, int64_t a,
                                  int64_t b) {
  int nf = ls.count;
  int64_t ca = a / nf, cb = b / nf;
  int fa = (int)(a % nf), fb = (int)(b % nf);
  for (int p = 0; p < ls.size; ++p)
    if (cells[ca * nv + ls.elem[fa][p]] != cells[cb * nv + ls.elem[fb][p]]) return false;
  return true;
}

__global__ void k_heads(const int32_t* cells, int nv, LocalSubsets ls, const uint32_t* perm,
                        int64_t n, int32_t* head) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  head[i] = (i == 0) ? 1 : (same_tuple(cells, nv, ls, perm[i], perm[i - 1]) ? 0 : 1);
}

// kidx = inclusive_scan(head) - 1; faces[slot] = kidx; rep[kidx] = slot at heads
__global__ void k_number(const uint32_t* perm, const int32_t* head, const int32_t* scan, int64_t n,
                         int32_t* faces, int64_t* rep) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t kidx = scan[i] - 1;
  faces[perm[i]] = kidx;
  if (head[i]) rep[kidx] = perm[i];
}

__global__ void k_compact_cells(const int32_t* cells_in, int nv, const uint32_t* perm,
                                const int32_t* head, const int32_t* scan, int64_t n,
                                int32_t* cells_out) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n || !head[i]) return;
  int64_t dst = scan[i] - 1;
  for (int p = 0; p < nv; ++p) cells_out[dst * nv + p] = cells_in[(int64_t)perm[i] * nv + p];
}

__global__ void k_count(const int32_t* idx, int64_t n, int32_t* cnt) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&cnt[idx[i]], 1);
}

// For every (cell, local facet) lying in exactly one cell, flag the facet and
// all its faces in every grade below (complex.rs:140-149).
struct FaceTables {
  int nsub[5];
  uint8_t mask[5][10];  // local subset masks per grade
};

__global__ void k_boundary(const int32_t* const* faces, int d, int64_t ncells, FaceTables ft,
                           const int32_t* facet_mult, uint8_t* const* bnd) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int nfac = d + 1;
  if (t >= ncells * nfac) return;
  int64_t cell = t / nfac;
  int f = (int)(t % nfac);
  int32_t g = faces[d - 1][cell * nfac + f];
  if (facet_mult[g] != 1) return;
  uint8_t fm = ft.mask[d - 1][f];
  for (int k = 0; k < d; ++k)
    for (int s = 0; s < ft.nsub[k]; ++s)
      if ((ft.mask[k][s] & ~fm) == 0) bnd[k][faces[k][cell * ft.nsub[k] + s]] = 1;
}

__global__ void k_interior_index(const uint8_t* bnd, const int32_t* scan, int64_t n, int32_t* out) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = bnd[i] ? -1 : scan[i] - 1;
}

template <class T>
void inclusive_scan(Ctx* ctx, const T* in, T* out, int64_t n) {
  size_t bytes = 0;
  FEEC_HIP(rocprim::inclusive_scan(nullptr, bytes, in, out, (size_t)n, rocprim::plus<T>(),
                                   ctx->stream));
  void* tmp = ctx->ensure_temp(bytes);
  FEEC_HIP(rocprim::inclusive_scan(tmp, bytes, in, out, (size_t)n, rocprim::plus<T>(),
                                   ctx->stream));
}

void sort_pairs(Ctx* ctx, uint64_t* kin, uint64_t* kout, uint32_t* vin, uint32_t* vout, int64_t n,
                int end_bit) {
  size_t bytes = 0;
  FEEC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (size_t)n, 0,
                                     (unsigned)end_bit, ctx->stream));
  void* tmp = ctx->ensure_temp(bytes);
  FEEC_HIP(rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout, (size_t)n, 0,
                                     (unsigned)end_bit, ctx->stream));
}

// Sort the n = ncells * ls.count tuples colex; returns the permutation (in
// `perm`) and run heads / inclusive scan of heads.
int64_t sort_tuples(Ctx* ctx, const int32_t* cells, int nv, const LocalSubsets& ls, int64_t n,
                    int64_t nverts, DBuf<uint32_t>& perm, DBuf<int32_t>& head,
                    DBuf<int32_t>& scan) {
  Packing pk = packing(nverts, ls.size);
  DBuf<uint64_t> k0(ctx, n), k1(ctx, n);
  DBuf<uint32_t> p1(ctx, n);
  perm.alloc(ctx, n);
  k_iota<<<grid_for(n), TPB, 0, ctx->stream>>>(perm.p, n);
  for (int w = 0; w < pk.words; ++w) {
    int nvert_w = std::min(pk.per_word, ls.size - w * pk.per_word);
    k_gather_words<<<grid_for(n), TPB, 0, ctx->stream>>>(cells, nv, ls, perm.p, n, w, pk.bits,
                                                          pk.per_word, k0.p);
    sort_pairs(ctx, k0.p, k1.p, perm.p, p1.p, n, pk.bits * nvert_w);
    std::swap(perm, p1);
  }
  head.alloc(ctx, n);
  scan.alloc(ctx, n);
  k_heads<<<grid_for(n), TPB, 0, ctx->stream>>>(cells, nv, ls, perm.p, n, head.p);
  inclusive_scan(ctx, head.p, scan.p, n);
  int32_t last = 0;
  FEEC_HIP(hipMemcpyAsync(&last, scan.p + (n - 1), sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
  ctx->sync();
  return last;
}

__global__ void k_coboundary(const int32_t* cells, int nv, LocalSubsets co, const int64_t* rep,
                             int64_t nrows, int k, const int32_t* faces_k, int nfk,
                             const uint8_t* face_of_deletion /* [co.count][k+2] */,
                             int64_t* row_ptr, int32_t* col, double* val) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r > nrows) return;
  int m = k + 2;
  row_ptr[r] = r * m;
  if (r == nrows) return;
  int64_t slot = (rep != nullptr) ? rep[r] : r * co.count;  // top grade: the cell itself
  int64_t cell = slot / co.count;
  int f = (int)(slot % co.count);
  int32_t c[6];
  double s[6];
  for (int p = 0; p < m; ++p) {
    c[p] = faces_k[cell * nfk + face_of_deletion[f * m + p]];
    s[p] = (p & 1) ? -1.0 : 1.0;
  }
  // columns ascending within the row
  for (int i = 1; i < m; ++i) {
    int32_t kc = c[i];
    double ks = s[i];
    int j = i - 1;
    while (j >= 0 && c[j] > kc) {
      c[j + 1] = c[j];
      s[j + 1] = s[j];
      --j;
    }
    c[j + 1] = kc;
    s[j + 1] = ks;
  }
  for (int p = 0; p < m; ++p) {
    col[r * m + p] = c[p];
    val[r * m + p] = s[p];
  }
}

__global__ void k_lengths(const int32_t* cells, int nv, LocalSubsets edges, const int64_t* rep,
                          int64_t ne, const double* coords, int d, int minkowski, double* out) {
  // Each square rounded on its own, then summed in axis order, with no fused
  // multiply-add: the squared lengths are bit-identical to the reference's
  // (Rust never contracts a*b+c), so exact-zero element entries of
  // exactly-representable cells come out the same.
#pragma clang fp contract(off)
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int64_t cell, vi, vj;
  if (rep) {
    cell = rep[e] / edges.count;
    int f = (int)(rep[e] % edges.count);
    vi = cells[cell * nv + edges.elem[f][0]];
    vj = cells[cell * nv + edges.elem[f][1]];
  } else {  // d == 1: the edges are the cells
    vi = cells[e * nv];
    vj = cells[e * nv + 1];
  }
  double acc = 0.0;
  for (int a = 0; a < d; ++a) {
    double t = coords[vj * d + a] - coords[vi * d + a];
    double sq = t * t;
    acc = (minkowski && a == 0) ? acc - sq : acc + sq;
  }
  out[e] = acc;
}

}  // namespace

void mesh_build(Mesh& m, int d, int64_t nverts, int64_t ncells, const int64_t* cells_host) {
  FEEC_REQUIRE(d >= 1 && d <= 4, FEEC_E_ARG, "dimension must be 1..4");
  FEEC_REQUIRE(ncells > 0 && nverts > 0, FEEC_E_ARG, "empty mesh");
  FEEC_REQUIRE(nverts < (int64_t(1) << 31), FEEC_E_ARG, "too many vertices");
  const int nv = d + 1;
  Ctx* ctx = m.ctx;
  m.d = d;
  // Host: sort each cell's vertices, validate (complex.rs:304-314).
  std::vector<int32_t> cells((size_t)ncells * nv);
  std::vector<uint8_t> used((size_t)nverts, 0);
  for (int64_t c = 0; c < ncells; ++c) {
    int64_t tmp[5];
    for (int p = 0; p < nv; ++p) {
      tmp[p] = cells_host[c * nv + p];
      FEEC_REQUIRE(tmp[p] >= 0 && tmp[p] < nverts, FEEC_E_TOPOLOGY, "cell vertex out of range");
    }
    std::sort(tmp, tmp + nv);
    for (int p = 0; p < nv; ++p) {
      FEEC_REQUIRE(p == 0 || tmp[p] != tmp[p - 1], FEEC_E_TOPOLOGY, "repeated vertex in a cell");
      cells[c * nv + p] = (int32_t)tmp[p];
      used[tmp[p]] = 1;
    }
  }
  for (int64_t v = 0; v < nverts; ++v)
    FEEC_REQUIRE(used[v], FEEC_E_TOPOLOGY,
                 "mesh vertices must be contiguous and fully used (no dangling vertices)");

  DBuf<int32_t> craw(ctx, cells.size());
  FEEC_HIP(hipMemcpyAsync(craw.p, cells.data(), cells.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice, ctx->stream));

  // Cells in colex order, deduplicated (skeleton.rs:60-62).
  LocalSubsets whole = local_subsets(d, d);
  {
    DBuf<uint32_t> perm;
    DBuf<int32_t> head, scan;
    int64_t nc = sort_tuples(ctx, craw.p, nv, whole, ncells, nverts, perm, head, scan);
    m.n[d] = nc;
    m.faces[0].alloc(ctx, (size_t)nc * nv);
    k_compact_cells<<<grid_for(ncells), TPB, 0, ctx->stream>>>(craw.p, nv, perm.p, head.p, scan.p,
                                                                ncells, m.faces[0].p);
  }
  craw.reset();
  const int64_t nc = m.n[d];
  m.n[0] = nverts;
  const int32_t* cs = m.faces[0].p;
  for (int k = 1; k < d; ++k) {
    LocalSubsets ls = local_subsets(d, k);
    int64_t n = nc * ls.count;
    FEEC_REQUIRE(n < (int64_t(1) << 32), FEEC_E_ARG, "mesh too large for 32-bit face slots");
    DBuf<uint32_t> perm;
    DBuf<int32_t> head, scan;
    m.n[k] = sort_tuples(ctx, cs, nv, ls, n, nverts, perm, head, scan);
    m.faces[k].alloc(ctx, (size_t)n);
    m.rep[k].alloc(ctx, (size_t)m.n[k]);
    k_number<<<grid_for(n), TPB, 0, ctx->stream>>>(perm.p, head.p, scan.p, n, m.faces[k].p,
                                                   m.rep[k].p);
  }

  // The cells are their own top-grade faces.
  m.faces[d].alloc(ctx, (size_t)nc);
  k_iota<<<grid_for(nc), TPB, 0, ctx->stream>>>(m.faces[d].p, nc);

  // Boundary flags.
  FaceTables ft{};
  for (int k = 0; k <= d; ++k) {
    int cnt = 0;
    for (unsigned msk = 0; msk < (1u << nv); ++msk)
      if (__builtin_popcount(msk) == k + 1) ft.mask[k][cnt++] = (uint8_t)msk;
    ft.nsub[k] = cnt;
  }
  for (int k = 0; k <= d; ++k) {
    m.bnd[k].alloc(ctx, (size_t)m.n[k]);
    FEEC_HIP(hipMemsetAsync(m.bnd[k].p, 0, m.n[k], ctx->stream));
  }
  {
    DBuf<int32_t> mult(ctx, (size_t)m.n[d - 1]);
    FEEC_HIP(hipMemsetAsync(mult.p, 0, m.n[d - 1] * sizeof(int32_t), ctx->stream));
    int64_t n = nc * (d + 1);
    k_count<<<grid_for(n), TPB, 0, ctx->stream>>>(m.faces[d - 1].p, n, mult.p);
    DBuf<const int32_t*> fptr(ctx, 5);
    DBuf<uint8_t*> bptr(ctx, 5);
    const int32_t* fh[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    uint8_t* bh[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    for (int k = 0; k < d; ++k) {
      fh[k] = m.faces[k].p;
      bh[k] = m.bnd[k].p;
    }
    FEEC_HIP(hipMemcpyAsync(fptr.p, fh, sizeof(fh), hipMemcpyHostToDevice, ctx->stream));
    FEEC_HIP(hipMemcpyAsync(bptr.p, bh, sizeof(bh), hipMemcpyHostToDevice, ctx->stream));
    k_boundary<<<grid_for(n), TPB, 0, ctx->stream>>>(fptr.p, d, nc, ft, mult.p, bptr.p);
    // pseudomanifold: a facet lies in one or two cells (complex.rs:278-290)
    std::vector<int32_t> hm((size_t)m.n[d - 1]);
    FEEC_HIP(hipMemcpyAsync(hm.data(), mult.p, hm.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                            ctx->stream));
    ctx->sync();
    for (int32_t c : hm)
      FEEC_REQUIRE(c == 1 || c == 2, FEEC_E_TOPOLOGY, "a facet must lie in one or two cells");
  }
  ctx->sync();
}

void Mesh::ensure_interior() {
  std::lock_guard<std::mutex> lock(mu);
  if (have_interior) return;
  for (int k = 0; k <= d; ++k) {
    int64_t nk = n[k];
    DBuf<int32_t> flag(ctx, nk), scan(ctx, nk);
    interior[k].alloc(ctx, nk);
    // interior flag = 1 - bnd
    std::vector<uint8_t> hb(nk);
    FEEC_HIP(hipMemcpyAsync(hb.data(), bnd[k].p, nk, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    std::vector<int32_t> hi(nk);
    int32_t cnt = 0;
    for (int64_t i = 0; i < nk; ++i) hi[i] = hb[i] ? -1 : cnt++;
    n_interior[k] = cnt;
    FEEC_HIP(hipMemcpyAsync(interior[k].p, hi.data(), nk * sizeof(int32_t), hipMemcpyHostToDevice,
                            ctx->stream));
    ctx->sync();
  }
  have_interior = true;
}

// Coherent orientation by propagation across shared facets: the two cells of an
// interior facet induce opposite orientations on it.
// Connected components of the vertex graph in which every cell joins its
// vertices (union-find over the cells' vertex lists).
void Mesh::ensure_components() {
  std::lock_guard<std::mutex> lock(mu);
  if (components >= 0) return;
  const int64_t nc = n[d], nv = n[0];
  const int nvert = d + 1;
  std::vector<int32_t> cv((size_t)nc * nvert);
  FEEC_HIP(hipMemcpyAsync(cv.data(), faces[0].p, cv.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
  ctx->sync();
  std::vector<int64_t> parent(nv);
  for (int64_t v = 0; v < nv; ++v) parent[v] = v;
  auto find = [&](int64_t v) {
    while (parent[v] != v) {
      parent[v] = parent[parent[v]];
      v = parent[v];
    }
    return v;
  };
  int64_t comps = nv;
  for (int64_t c = 0; c < nc; ++c) {
    const int64_t r0 = find(cv[c * nvert]);
    for (int i = 1; i < nvert; ++i) {
      const int64_t r = find(cv[c * nvert + i]);
      if (r != r0) {
        parent[r] = r0;
        --comps;
      }
    }
  }
  components = comps;
}

void Mesh::ensure_orientation() {
  std::lock_guard<std::mutex> lock(mu);
  if (have_orientation) return;
  int64_t nc = n[d], nf = n[d - 1];
  int nfac = d + 1;
  std::vector<int32_t> fc((size_t)nc * nfac);
  FEEC_HIP(hipMemcpyAsync(fc.data(), faces[d - 1].p, fc.size() * sizeof(int32_t),
                          hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  // local facet f (colex d-subsets of d+1 positions) deletes position d - f
  std::vector<int64_t> adj((size_t)nf * 2, -1);
  std::vector<int8_t> sgn((size_t)nf * 2, 0);
  for (int64_t c = 0; c < nc; ++c)
    for (int f = 0; f < nfac; ++f) {
      int64_t g = fc[c * nfac + f];
      int p = d - f;
      int slot = adj[g * 2] < 0 ? 0 : 1;
      adj[g * 2 + slot] = c;
      sgn[g * 2 + slot] = (p & 1) ? -1 : 1;
    }
  orientation.assign(nc, 0);
  orientable = true;
  std::vector<int64_t> stack;
  for (int64_t root = 0; root < nc && orientable; ++root) {
    if (orientation[root]) continue;
    orientation[root] = 1;
    stack.push_back(root);
    while (!stack.empty() && orientable) {
      int64_t c = stack.back();
      stack.pop_back();
      for (int f = 0; f < nfac; ++f) {
        int64_t g = fc[c * nfac + f];
        if (adj[g * 2 + 1] < 0) continue;
        int me = adj[g * 2] == c ? 0 : 1;
        int64_t o = adj[g * 2 + 1 - me];
        int8_t want = (int8_t)(-orientation[c] * sgn[g * 2 + me] * sgn[g * 2 + 1 - me]);
        if (!orientation[o]) {
          orientation[o] = want;
          stack.push_back(o);
        } else if (orientation[o] != want) {
          orientable = false;
        }
      }
    }
  }
  have_orientation = true;
}

void mesh_coboundary(Mesh& m, int k, Csr& out) {
  Ctx* ctx = m.ctx;
  int d = m.d;
  out.ctx = ctx;
  FEEC_REQUIRE(k >= 0 && k < d, FEEC_E_ARG, "coboundary grade out of range");
  int64_t nrows = m.n[k + 1];
  int mm = k + 2;
  LocalSubsets co = local_subsets(d, k + 1), fa = local_subsets(d, k);
  // local face index of each deletion of each local (k+1)-face
  std::vector<uint8_t> del((size_t)co.count * mm);
  for (int s = 0; s < co.count; ++s) {
    unsigned cm = 0;
    for (int p = 0; p < mm; ++p) cm |= 1u << co.elem[s][p];
    for (int p = 0; p < mm; ++p) {
      unsigned fm = cm & ~(1u << co.elem[s][p]);
      for (int q = 0; q < fa.count; ++q) {
        unsigned qm = 0;
        for (int r = 0; r < k + 1; ++r) qm |= 1u << fa.elem[q][r];
        if (qm == fm) del[s * mm + p] = (uint8_t)q;
      }
    }
  }
  DBuf<uint8_t> ddel(ctx, del.size());
  FEEC_HIP(hipMemcpyAsync(ddel.p, del.data(), del.size(), hipMemcpyHostToDevice, ctx->stream));
  out.nrows = nrows;
  out.ncols = m.n[k];
  out.nnz = nrows * mm;
  out.row_ptr.alloc(ctx, nrows + 1);
  out.col.alloc(ctx, out.nnz);
  out.val.alloc(ctx, out.nnz);
  const int64_t* rep = (k + 1 < d) ? m.rep[k + 1].p : nullptr;
  const int32_t* fk = (k == 0) ? m.faces[0].p : m.faces[k].p;
  k_coboundary<<<grid_for(nrows + 1), TPB, 0, ctx->stream>>>(
      m.faces[0].p, d + 1, co, rep, nrows, k, fk, m.nfaces(k), ddel.p, out.row_ptr.p, out.col.p,
      out.val.p);
  ctx->sync();
}

void lengths_from_coords(Mesh& m, const double* coords_dev, int minkowski, double* out_dev) {
  LocalSubsets e = local_subsets(m.d, 1);
  const int64_t* rep = (m.d > 1) ? m.rep[1].p : nullptr;
  k_lengths<<<grid_for(m.n[1]), TPB, 0, m.ctx->stream>>>(m.faces[0].p, m.d + 1, e, rep, m.n[1],
                                                         coords_dev, m.d, minkowski, out_dev);
}

}  // namespace feec
