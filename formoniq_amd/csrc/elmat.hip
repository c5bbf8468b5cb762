// Per-cell Regge metric and Whitney element matrices, FP64, one thread per cell.
//
// Reference behaviour (luiswirth/formoniq, crates/):
//   SimplexLengthsSq::metric       regge/src/lengths/simplex.rs:308-326
//   Metric::{det, dual}            metric/src/lib.rs:179-190, 250-257
//   cell_volume                    regge/src/lib.rs:26-28
//   multiform_metric               metric/src/tensor.rs:127-131
//   HodgeMass::element             formoniq/src/operators.rs:84-94
//   WhitneyExpansion::pullback     derham/src/interpolate/form.rs:195-232
//   WhitneyPairing::element        formoniq/src/operators.rs:201-211
//
// Formulation. As the reference, H = P (Lambda^K g^-1) P^T with P the K-th
// compound of the barycentric differentials Dlambda (row 0 = -1, rows 1..d =
// I); P has entries 0 and +-1 in closed form. The pullback sums (K+1)^2 signed
// H entries per mass entry with weights 1 or 2 (the barycentric Gramian
// Q = (1 + delta)/((d+1)(d+2))), and the exterior derivative sides are +-1
// incidences, so every element block -- mass, dif_trial, dif_test, dif_both --
// is a fixed integer-coefficient linear map of the NU = n(n+1)/2 unique
// entries of the symmetric Lambda^K g^-1 (n = C(d,K), NU <= 21 for d <= 4),
// times vol (K!)^2/((d+1)(d+2)). The integer maps are built once on the host.
//
// Kernel shape: one wavefront per group of 64 cells, one lane per cell. The
// lane computes g, det g and g^-1 in registers, then for each grade K the
// unique K x K minors L of g^-1 in registers, and emits every block of that
// grade as a dense contraction out_e = scale * sum_u C[e][u] L[u]. C is
// wave-uniform: it is read with scalar loads and enters the FP64 FMAs as an
// SGPR operand, so the inner loop touches neither LDS nor vector memory.
// Element matrices are written structure-of-arrays,
// out[entry * ncells + cell], so each store instruction of a wavefront covers
// 64 consecutive doubles.
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "feec_internal.hpp"

namespace feec {

namespace {

constexpr int TPB = 256;

constexpr int popc(unsigned m) {
  int c = 0;
  while (m) {
    c += m & 1u;
    m >>= 1;
  }
  return c;
}

constexpr int colex_rank(unsigned m) {
  int r = 0, i = 0;
  for (int e = 0; e < 32; ++e)
    if (m & (1u << e)) {
      r += binom(e, i + 1);
      ++i;
    }
  return r;
}

constexpr double factorial(int n) { return n <= 1 ? 1.0 : n * factorial(n - 1); }

template <int K>
__device__ __forceinline__ double det(const double (&a)[K > 0 ? K : 1][K > 0 ? K : 1]) {
  if constexpr (K == 0) {
    return 1.0;
  } else if constexpr (K == 1) {
    return a[0][0];
  } else if constexpr (K == 2) {
    return a[0][0] * a[1][1] - a[0][1] * a[1][0];
  } else if constexpr (K == 3) {
    return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
           a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
           a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  } else {
    static_assert(K == 4, "K <= 4");
    // expansion along row 0 with 2x2 minors of rows 2,3 shared
    double s01 = a[2][0] * a[3][1] - a[2][1] * a[3][0];
    double s02 = a[2][0] * a[3][2] - a[2][2] * a[3][0];
    double s03 = a[2][0] * a[3][3] - a[2][3] * a[3][0];
    double s12 = a[2][1] * a[3][2] - a[2][2] * a[3][1];
    double s13 = a[2][1] * a[3][3] - a[2][3] * a[3][1];
    double s23 = a[2][2] * a[3][3] - a[2][3] * a[3][2];
    double c0 = a[1][1] * s23 - a[1][2] * s13 + a[1][3] * s12;
    double c1 = a[1][0] * s23 - a[1][2] * s03 + a[1][3] * s02;
    double c2 = a[1][0] * s13 - a[1][1] * s03 + a[1][3] * s01;
    double c3 = a[1][0] * s12 - a[1][1] * s02 + a[1][2] * s01;
    return a[0][0] * c0 - a[0][1] * c1 + a[0][2] * c2 - a[0][3] * c3;
  }
}

// Metric by polarization, det and inverse by the adjugate.
template <int D>
struct CellGeom {
  double g[D][D];
  double ginv[D][D];
  double det;
};

__device__ __forceinline__ constexpr int edge_index(int i, int j) { return i + j * (j - 1) / 2; }

template <int D>
__device__ __forceinline__ void cell_geometry(const double* s, CellGeom<D>& cg) {
#pragma unroll
  for (int i = 0; i < D; ++i) cg.g[i][i] = s[edge_index(0, i + 1)];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = i + 1; j < D; ++j) {
      double v = 0.5 * (s[edge_index(0, i + 1)] + s[edge_index(0, j + 1)] -
                        s[edge_index(i + 1, j + 1)]);
      cg.g[i][j] = v;
      cg.g[j][i] = v;
    }
  cg.det = det<D>(cg.g);
  const double inv = 1.0 / cg.det;
  if constexpr (D == 1) {
    cg.ginv[0][0] = inv;
  } else {
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        double m[D - 1][D - 1];
#pragma unroll
        for (int a = 0; a < D - 1; ++a)
#pragma unroll
          for (int b = 0; b < D - 1; ++b) m[a][b] = cg.g[a < i ? a : a + 1][b < j ? b : b + 1];
        double cof = det<D - 1>(m);
        cg.ginv[j][i] = (((i + j) & 1) ? -cof : cof) * inv;
      }
  }
}

// ---------------------------------------------------------------------------
// coefficient tables (host)
// ---------------------------------------------------------------------------
// Block kinds: 0 mass, 1 dif_trial, 2 dif_test, 3 dif_both, 4 diag(mass), 5 diag(dif_both)
constexpr int KIND_DIAG_MASS = 4, KIND_DIAG_BOTH = 5;

struct HostBlock {
  int K = 0, kind = 0, rows = 0, cols = 0;
  std::vector<int32_t> start;  // nout + 1
  std::vector<uint16_t> h;
  std::vector<int16_t> c;
};

static std::vector<unsigned> masks_of(int n, int size) {
  std::vector<unsigned> out;
  if (size < 0 || size > n) return out;
  for (unsigned m = 0; m < (1u << n); ++m)
    if (popc(m) == size) out.push_back(m);
  return out;
}
static int elem_of(unsigned m, int p) {
  for (int e = 0; e < 32; ++e)
    if (m & (1u << e)) {
      if (p == 0) return e;
      --p;
    }
  return -1;
}
// index of the symmetric pair (a, b), a <= b, row-major over the upper triangle
static int sym_index(int a, int b, int nb) {
  if (a > b) std::swap(a, b);
  return a * nb - a * (a - 1) / 2 + (b - a);
}

using LinH = std::map<int, long>;  // unique Lambda^K g^-1 index -> integer coefficient

// Entry (a, S) of P = Lambda^K(Dlambda), Dlambda row 0 = -1 and rows v >= 1 =
// e_{v-1}: a is a vertex subset (bit mask), S a subset of the d axes.
static int compound_dlambda(unsigned a, unsigned S) {
  if (!(a & 1u)) return S == (a >> 1) ? 1 : 0;
  const unsigned rest = (a & ~1u) >> 1;
  if ((S & rest) != rest || popc(S) != popc(a)) return 0;
  const unsigned s = S & ~rest;
  if (popc(s) != 1) return 0;
  return (popc(S & (s - 1u)) & 1) ? 1 : -1;
}

static HostBlock make_block(int D, int K, int kind) {
  const int nv = D + 1;
  auto blades = masks_of(nv, K);
  auto dofs = masks_of(nv, K + 1);
  auto axes = masks_of(D, K);
  const int nb = (int)blades.size(), nd = (int)dofs.size(), na = (int)axes.size();
  auto blade_idx = [&](unsigned m) {
    for (int i = 0; i < nb; ++i)
      if (blades[i] == m) return i;
    return -1;
  };
  // H[a][b] = sum_{S,T} P[a][S] P[b][T] Lambda[S][T] (operators.rs:76, 84-94)
  std::vector<LinH> Hl((size_t)nb * nb);
  for (int a = 0; a < nb; ++a)
    for (int b = 0; b < nb; ++b)
      for (int S = 0; S < na; ++S) {
        int pa = compound_dlambda(blades[a], axes[S]);
        if (!pa) continue;
        for (int T = 0; T < na; ++T) {
          int pb = compound_dlambda(blades[b], axes[T]);
          if (pb) Hl[(size_t)a * nb + b][sym_index(S, T, na)] += (long)pa * pb;
        }
      }
  // Whitney mass M[s][t] as a combination of Lambda entries (form.rs:195-232)
  std::vector<LinH> M((size_t)nd * nd);
  for (int i = 0; i < nd; ++i)
    for (int j = 0; j < nd; ++j) {
      LinH acc;
      for (int p = 0; p <= K; ++p)
        for (int q = 0; q <= K; ++q) {
          int va = elem_of(dofs[i], p), vb = elem_of(dofs[j], q);
          int a = blade_idx(dofs[i] & ~(1u << va)), b = blade_idx(dofs[j] & ~(1u << vb));
          long w = (((p + q) & 1) ? -1 : 1) * (va == vb ? 2 : 1);
          for (auto& kv : Hl[(size_t)a * nb + b]) acc[kv.first] += w * kv.second;
        }
      M[(size_t)i * nd + j] = acc;
    }
  // local boundary d_K: rows = K-subsets (faces), cols = dofs (simplex.rs:223-238)
  std::vector<int> bd((size_t)nb * nd, 0);
  if (K >= 1)
    for (int s = 0; s < nd; ++s)
      for (int p = 0; p <= K; ++p) {
        int f = blade_idx(dofs[s] & ~(1u << elem_of(dofs[s], p)));
        bd[(size_t)f * nd + s] = (p & 1) ? -1 : 1;
      }
  auto add = [](LinH& dst, const LinH& src, long w) {
    for (auto& kv : src) dst[kv.first] += w * kv.second;
  };
  HostBlock hb;
  hb.K = K;
  hb.kind = kind;
  std::vector<LinH> out;
  if (kind == 0 || kind == KIND_DIAG_MASS) {
    hb.rows = nd;
    hb.cols = nd;
    out = M;
  } else if (kind == 1) {  // M d^T
    hb.rows = nd;
    hb.cols = nb;
    out.assign((size_t)nd * nb, {});
    for (int s = 0; s < nd; ++s)
      for (int f = 0; f < nb; ++f)
        for (int t = 0; t < nd; ++t)
          if (bd[(size_t)f * nd + t]) add(out[(size_t)s * nb + f], M[(size_t)s * nd + t], bd[(size_t)f * nd + t]);
  } else if (kind == 2) {  // d M
    hb.rows = nb;
    hb.cols = nd;
    out.assign((size_t)nb * nd, {});
    for (int f = 0; f < nb; ++f)
      for (int t = 0; t < nd; ++t)
        for (int s = 0; s < nd; ++s)
          if (bd[(size_t)f * nd + s]) add(out[(size_t)f * nd + t], M[(size_t)s * nd + t], bd[(size_t)f * nd + s]);
  } else {  // d M d^T
    hb.rows = nb;
    hb.cols = nb;
    out.assign((size_t)nb * nb, {});
    for (int f = 0; f < nb; ++f)
      for (int g = 0; g < nb; ++g)
        for (int s = 0; s < nd; ++s)
          for (int t = 0; t < nd; ++t) {
            long w = (long)bd[(size_t)f * nd + s] * bd[(size_t)g * nd + t];
            if (w) add(out[(size_t)f * nb + g], M[(size_t)s * nd + t], w);
          }
  }
  std::vector<int> pick;
  if (kind == KIND_DIAG_MASS || kind == KIND_DIAG_BOTH) {
    for (int i = 0; i < hb.rows; ++i) pick.push_back(i * hb.cols + i);
    hb.cols = 1;
  } else {
    for (int e = 0; e < hb.rows * hb.cols; ++e) pick.push_back(e);
  }
  hb.start.push_back(0);
  for (int e : pick) {
    for (auto& kv : out[e])
      if (kv.second != 0) {
        hb.h.push_back((uint16_t)kv.first);
        hb.c.push_back((int16_t)kv.second);
      }
    hb.start.push_back((int32_t)hb.h.size());
  }
  return hb;
}

// Device-resident coefficient tables of one dimension: per (grade, kind)
// block, a dense [nout][NU] table of integer coefficients (stored as double)
// over the NU unique minors of that grade.
struct DevTables {
  DBuf<double> coef;
  int nU[5] = {0, 0, 0, 0, 0};
  int block_coef_off[5][6];
  int block_nout[5][6];
};

static std::mutex g_tab_mu;
static std::map<std::pair<Ctx*, int>, std::unique_ptr<DevTables>> g_tabs;

}  // namespace

void elmat_release_tables(Ctx* ctx) {
  std::lock_guard<std::mutex> lock(g_tab_mu);
  for (auto it = g_tabs.begin(); it != g_tabs.end();)
    it = (it->first.first == ctx) ? g_tabs.erase(it) : std::next(it);
}

namespace {

static DevTables& tables_for(Ctx* ctx, int D) {
  std::lock_guard<std::mutex> lock(g_tab_mu);
  auto key = std::make_pair(ctx, D);
  auto it = g_tabs.find(key);
  if (it != g_tabs.end()) return *it->second;
  auto T = std::make_unique<DevTables>();
  std::vector<double> coef;
  const int nv = D + 1;
  for (int K = 0; K <= D; ++K) {
    const int na = binom(D, K);
    const int nu = na * (na + 1) / 2;
    T->nU[K] = nu;
    for (int kind = 0; kind < 6; ++kind) {
      T->block_nout[K][kind] = 0;
      T->block_coef_off[K][kind] = (int)coef.size();
      if (K == 0 && kind != 0 && kind != KIND_DIAG_MASS) continue;
      HostBlock hb = make_block(D, K, kind);
      const int nout = (int)hb.start.size() - 1;
      T->block_nout[K][kind] = nout;
      size_t base = coef.size();
      coef.resize(base + (size_t)nout * nu, 0.0);
      for (int e = 0; e < nout; ++e)
        for (int t = hb.start[e]; t < hb.start[e + 1]; ++t)
          coef[base + (size_t)e * nu + hb.h[t]] += (double)hb.c[t];
    }
  }
  T->coef.alloc(ctx, std::max<size_t>(coef.size(), 1));
  copy_to_device(ctx, T->coef.p, coef.data(), coef.size() * sizeof(double));
  auto& slot = g_tabs[key];
  slot = std::move(T);
  return *slot;
}

// A launch request: the grades whose minors are needed and the blocks emitted.
struct JobBlock {
  int K;
  int nout;
  int coef_off;  // into the dense coefficient table of the dimension
  double* out;   // SoA [nout][ncells]
};
struct Job {
  int ngrades;
  int grades[5];
  int nblocks;
  JobBlock b[12];
};

constexpr int WAVE = 64;

// The K-subsets of {0..N-1} in colex order (= increasing bit masks).
template <int N, int K>
struct Blades {
  static constexpr int count = binom(N, K);
  int e[count > 0 ? count : 1][K > 0 ? K : 1];
  constexpr Blades() : e() {
    int i = 0;
    for (unsigned m = 0; m < (1u << N); ++m) {
      int c = 0;
      for (int b = 0; b < N; ++b) c += (m >> b) & 1u;
      if (c != K) continue;
      int p = 0;
      for (int b = 0; b < N; ++b)
        if ((m >> b) & 1u) e[i][p++] = b;
      ++i;
    }
  }
};

// Coefficients are read through the constant address space, so the compiler
// issues wave-uniform scalar loads (s_load_dwordx16) and the FMAs take them as
// SGPR operands.
typedef const __attribute__((address_space(4))) double* ConstDouble;

// Every block of grade K: the unique K x K minors L of g^-1 in registers,
// then out_e = scale * sum_u C[e][u] L[u] with the dense integer table C.
template <int D, int K>
__device__ __forceinline__ void grade_blocks(const double (&gi)[D][D], double vol, const Job& job,
                                             const double* __restrict__ coef, int64_t nc,
                                             int64_t cell, bool active) {
  constexpr Blades<D, K> B{};
  constexpr int NB = Blades<D, K>::count;
  constexpr int NU = NB * (NB + 1) / 2;
  double H[NU];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = a; b < NB; ++b) {
      double m[K > 0 ? K : 1][K > 0 ? K : 1];
#pragma unroll
      for (int i = 0; i < K; ++i)
#pragma unroll
        for (int j = 0; j < K; ++j) m[i][j] = gi[B.e[a][i]][B.e[b][j]];
      H[a * NB - a * (a - 1) / 2 + (b - a)] = det<K>(m);
    }
  constexpr double kf = factorial(K);
  const double scale = vol * (kf * kf / double((D + 1) * (D + 2)));
  for (int bi = 0; bi < job.nblocks; ++bi) {
    if (job.b[bi].K != K) continue;
    ConstDouble c = (ConstDouble)(coef + job.b[bi].coef_off);
    double* __restrict__ out = job.b[bi].out;
    const int nout = job.b[bi].nout;
    // small tables: two outputs per pass (independent FMA chains on the same
    // H operands); large ones keep one row of coefficients in SGPRs
    constexpr int PAIR = NU <= 28 ? 2 : 1;
    int e = 0;
    for (; PAIR == 2 && e + 1 < nout; e += 2, c += 2 * NU) {
      double a0 = 0.0, a1 = 0.0;
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        a0 += c[u] * H[u];
        a1 += c[NU + u] * H[u];
      }
      if (active) {
        out[(int64_t)e * nc + cell] = scale * a0;
        out[(int64_t)(e + 1) * nc + cell] = scale * a1;
      }
    }
    for (; e < nout; ++e, c += NU) {
      double a0 = 0.0;
#pragma unroll
      for (int u = 0; u < NU; ++u) a0 += c[u] * H[u];
      if (active) out[(int64_t)e * nc + cell] = scale * a0;
    }
  }
}

// One wavefront, 64 consecutive cells starting at cell0, one lane per cell.
template <int D>
__device__ __forceinline__ void elmat_group(const int32_t* __restrict__ edges,
                                            const double* __restrict__ len, int64_t nc,
                                            int64_t cell0, const Job& job,
                                            const double* __restrict__ coef, int* degenerate) {
  constexpr int NE = binom(D + 1, 2);
  const int64_t cell = cell0 + threadIdx.x;
  const bool active = cell < nc;
  double s[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) s[e] = active ? len[edges[cell * NE + e]] : ((e < D) ? 1.0 : 2.0);
  CellGeom<D> cg;
  cell_geometry<D>(s, cg);
  if (active && cg.det == 0.0) atomicOr(degenerate, 1);
  const double vol = (1.0 / factorial(D)) * sqrt(fabs(cg.det));
  for (int gk = 0; gk < job.ngrades; ++gk) {
    switch (job.grades[gk]) {
      case 0: grade_blocks<D, 0>(cg.ginv, vol, job, coef, nc, cell, active); break;
      case 1: grade_blocks<D, 1>(cg.ginv, vol, job, coef, nc, cell, active); break;
      case 2: if constexpr (D >= 2) grade_blocks<D, 2>(cg.ginv, vol, job, coef, nc, cell, active); break;
      case 3: if constexpr (D >= 3) grade_blocks<D, 3>(cg.ginv, vol, job, coef, nc, cell, active); break;
      case 4: if constexpr (D >= 4) grade_blocks<D, 4>(cg.ginv, vol, job, coef, nc, cell, active); break;
      default: break;
    }
  }
}

template <int D>
__global__ void __launch_bounds__(WAVE) k_elmat(const int32_t* __restrict__ edges,
                                                const double* __restrict__ len, int64_t nc, Job job,
                                                const double* __restrict__ coef, int* degenerate) {
  elmat_group<D>(edges, len, nc, blockIdx.x * (int64_t)WAVE, job, coef, degenerate);
}

// Many meshes of mixed dimension in one persistent launch: every wavefront
// takes (mesh, 64-cell range) items from a global counter until the queue is
// empty, so the grid drains whatever the mix of sizes.
struct BatchMesh {
  int D;
  int64_t nc;
  const int32_t* edges;
  const double* len;
  int64_t item0;  // first work item of this mesh
  Job job;
};

__global__ void __launch_bounds__(WAVE) k_elmat_batched(const BatchMesh* __restrict__ meshes,
                                                        int nmeshes, int64_t nitems,
                                                        unsigned long long* counter,
                                                        const double* __restrict__ coef2,
                                                        const double* __restrict__ coef3,
                                                        const double* __restrict__ coef4,
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
      case 2: elmat_group<2>(m.edges, m.len, m.nc, cell0, m.job, coef2, degenerate); break;
      case 3: elmat_group<3>(m.edges, m.len, m.nc, cell0, m.job, coef3, degenerate); break;
      case 4: elmat_group<4>(m.edges, m.len, m.nc, cell0, m.job, coef4, degenerate); break;
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

// Launch k_elmat for a job on the mesh of g.
void run_job(Geom& g, Job& job, bool check) {
  Mesh& m = *g.mesh;
  Ctx* ctx = m.ctx;
  const int D = m.d;
  int64_t nc = m.n[D];
  DevTables& T = tables_for(ctx, D);
  int* deg = ctx->dflag;
  FEEC_HIP(hipMemsetAsync(deg, 0, sizeof(int), ctx->stream));
  unsigned grid = (unsigned)std::max<int64_t>(1, (nc + WAVE - 1) / WAVE);
  const int32_t* edges = m.faces[1].p;
  const double* cf = T.coef.p;
  switch (D) {
    case 1: k_elmat<1><<<grid, WAVE, 0, ctx->stream>>>(edges, g.lengths.p, nc, job, cf, deg); break;
    case 2: k_elmat<2><<<grid, WAVE, 0, ctx->stream>>>(edges, g.lengths.p, nc, job, cf, deg); break;
    case 3: k_elmat<3><<<grid, WAVE, 0, ctx->stream>>>(edges, g.lengths.p, nc, job, cf, deg); break;
    case 4: k_elmat<4><<<grid, WAVE, 0, ctx->stream>>>(edges, g.lengths.p, nc, job, cf, deg); break;
  }
  FEEC_HIP(hipGetLastError());
  if (check) check_degenerate(ctx, deg);
}

Job kkt_job(DevTables& T, int d, int k, double* const* outs);

// Autotuned persistent grid (in wavefronts) per device, chosen by timing.
std::mutex g_tune_mu;
std::map<int, unsigned> g_tuned_waves;

void job_add_grade(Job& job, int K) {
  for (int i = 0; i < job.ngrades; ++i)
    if (job.grades[i] == K) return;
  job.grades[job.ngrades++] = K;
}

void job_add_block(Job& job, DevTables& T, int K, int kind, double* out) {
  FEEC_REQUIRE(job.nblocks < 12, FEEC_E_ARG, "too many element blocks in one launch");
  job_add_grade(job, K);
  JobBlock& b = job.b[job.nblocks++];
  b.K = K;
  b.nout = T.block_nout[K][kind];
  b.coef_off = T.block_coef_off[K][kind];
  b.out = out;
}

}  // namespace

void element_matrices_soa(Geom& g, int kind, int grade, double* out_dev) {
  Mesh& m = *g.mesh;
  DevTables& T = tables_for(m.ctx, m.d);
  Job job{};
  job_add_block(job, T, grade, kind, out_dev);
  run_job(g, job, true);
}

namespace {
Job kkt_job(DevTables& T, int d, int k, double* const* outs) {
  Job job{};
  if (k >= 1) {
    job_add_block(job, T, k - 1, 0, outs[0]);
    job_add_block(job, T, k, 2, outs[1]);
    job_add_block(job, T, k, KIND_DIAG_BOTH, outs[3]);
  }
  job_add_block(job, T, k, KIND_DIAG_MASS, outs[4]);
  if (k + 1 <= d) job_add_block(job, T, k + 1, 3, outs[2]);
  // the two hdif_gram blocks of the preconditioner need M_k and dif_both(k) in full
  job_add_block(job, T, k, 0, outs[5]);
  if (k >= 1) job_add_block(job, T, k, 3, outs[6]);
  return job;
}
}  // namespace

// outs: M_{k-1}, dif_test(k), dif_both(k+1), diag(dif_both(k)), diag(M_k), M_k, dif_both(k)
void kkt_element_blocks(Geom& g, int grade, double* const* outs, bool check) {
  Mesh& m = *g.mesh;
  DevTables& T = tables_for(m.ctx, m.d);
  Job job = kkt_job(T, m.d, grade, outs);
  run_job(g, job, check);
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
    DevTables& T = tables_for(ctx, m.d);
    BatchMesh& b = host[i];
    b.D = m.d;
    b.nc = m.n[m.d];
    b.edges = m.faces[1].p;
    b.len = reqs[i].g->lengths.p;
    b.item0 = items;
    b.job = kkt_job(T, m.d, reqs[i].grade, reqs[i].outs);
    items += (b.nc + WAVE - 1) / WAVE;
  }
  DevTables& T2 = tables_for(ctx, 2);
  DevTables& T3 = tables_for(ctx, 3);
  DevTables& T4 = tables_for(ctx, 4);
  DBuf<BatchMesh> dm(ctx, host.size());
  FEEC_HIP(hipMemcpyAsync(dm.p, host.data(), host.size() * sizeof(BatchMesh), hipMemcpyHostToDevice,
                          ctx->stream));
  FEEC_HIP(hipMemsetAsync(ctx->dflag, 0, sizeof(int), ctx->stream));
  auto launch = [&](unsigned waves) {
    FEEC_HIP(hipMemsetAsync(ctx->dcount, 0, sizeof(unsigned long long), ctx->stream));
    k_elmat_batched<<<waves, WAVE, 0, ctx->stream>>>(dm.p, (int)host.size(), items, ctx->dcount,
                                                     T2.coef.p, T3.coef.p, T4.coef.p, ctx->dflag);
    FEEC_HIP(hipGetLastError());
  };
  unsigned waves = 0;
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tuned_waves.find(ctx->device);
    if (it != g_tuned_waves.end()) waves = it->second;
  }
  hipEvent_t e0, e1;
  FEEC_HIP(hipEventCreate(&e0));
  FEEC_HIP(hipEventCreate(&e1));
  float best_ms = 0.f;
  if (waves == 0) {
    const unsigned cand[] = {1024, 2048, 4096, 8192, 16384};
    float best = 1e30f;
    for (unsigned c : cand) {
      FEEC_HIP(hipEventRecord(e0, ctx->stream));
      launch(c);
      FEEC_HIP(hipEventRecord(e1, ctx->stream));
      FEEC_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      FEEC_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        waves = c;
      }
    }
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tuned_waves[ctx->device] = waves;
  }
  FEEC_HIP(hipEventRecord(e0, ctx->stream));
  launch(waves);
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
  const int d = m.d;
  const int64_t nc = m.n[d];
  DevTables& T = tables_for(m.ctx, d);
  Job job{};
  int64_t off = 0;
  for (int k = 0; k <= d; ++k) {
    job_add_block(job, T, k, 0, out_dev + off * nc);
    off += T.block_nout[k][0];
    if (k >= 1) {
      job_add_block(job, T, k, 1, out_dev + off * nc);
      off += T.block_nout[k][1];
    }
  }
  run_job(g, job, check);
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
