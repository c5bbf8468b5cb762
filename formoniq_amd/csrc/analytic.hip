// Analytic data on the device: Grundmann-Moller quadrature, SourceForm loads,
// the Whitney interpolant at quadrature nodes, the L2 error and projection,
// and the de Rham map.
//
// Reference behaviour (luiswirth/formoniq, crates/):
//   SimplexQuadRule::grundmann_moeller   simplicial/src/atlas/quadrature.rs:37-83
//   SimplexQuadRule::integrate_unit      simplicial/src/atlas/quadrature.rs:100-116
//   CellQuadrature::integrate            formoniq/src/operators.rs:240-256
//   SourceForm::element                  formoniq/src/operators.rs:599-635
//   LinearForm assembly (cell order)     formoniq/src/galerkin.rs:205-312
//   WhitneyLsf::at_bary                  derham/src/interpolate/form.rs:109-114
//   WhitneyInterpolant::eval             derham/src/interpolate/interpolant.rs:55-62
//   Pullback::at, bary2global            derham/src/section.rs:182-195,
//                                        simplicial/src/atlas/simplex_coords.rs:196-205
//   fe_l2_error, l2_projection           formoniq/src/fe.rs:50-123
//   derham_map, integrate_face           derham/src/project.rs:55-96
//
// A field arrives as its values at quadrature nodes, [cell][node][component]
// (or [simplex][node][component] for the de Rham map), each k-form value on
// the colex basis dx^I. Kernels run one lane per cell (or per k-simplex): the
// lane computes its cell's metric from the squared edge lengths exactly as the
// element kernels do, Lambda^k g^-1, and, for an ambient-frame field, the
// pullback Lambda^k(J) of the cell's affine map; the rule's nodes and weights
// are read from a small device table by every lane (uniform addresses). A
// cell's element load is written structure-of-arrays [slot][cell] and summed
// per face through a cached COO plan whose contributions are in cell order,
// so each entry of the load vector is summed over its cells in the order the
// reference's assembly adds them.
#include <algorithm>
#include <cmath>
#include <vector>

#include "cell_geometry.hpp"
#include "feec_internal.hpp"

// The node loops carry "#pragma unroll" for the fixed-rule instances; in the
// generic instance their trip count is a run-time value and stays a loop.
#pragma clang diagnostic ignored "-Wpass-failed"

namespace feec {

namespace {

constexpr int TPB = 256;
inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + TPB - 1) / TPB); }

constexpr int kMaxRuleDegree = 7;
// key of the cached face-summation plans in Mesh::plans (disjoint from the
// assembly plans, whose first key is a feec_bc)
constexpr int kLoadPlanTag = 16;
// the reference's StopCriterion::rtol default budget (iterative/src/lib.rs:202-207)
constexpr int64_t kProjectionMaxIters = 10000;
constexpr double kProjectionRtol = 1e-12;

// ---------------------------------------------------------------------------
// Grundmann-Moller rules (host)
// ---------------------------------------------------------------------------
struct Rule {
  int dim = 0, np = 0;
  std::vector<double> bary;  // [np][dim + 1]
  std::vector<double> w;     // [np], summing to one
};

// weakly increasing words of length len over nsym symbols, in colex order
// (compare from the last letter): Composition::all's order
void colex_words(int nsym, int len, std::vector<std::vector<int>>& out) {
  out.clear();
  if (len == 0) {
    out.push_back({});
    return;
  }
  std::vector<int> w(len, 0);
  for (;;) {
    out.push_back(w);
    // colex successor of a weakly increasing word: raise the first letter
    // that may grow (w[i] < w[i+1], or the last below nsym-1), reset the
    // letters before it to 0
    int i = 0;
    while (i < len && !((i + 1 < len) ? (w[i] < w[i + 1]) : (w[i] < nsym - 1))) ++i;
    if (i == len) break;
    ++w[i];
    for (int j = 0; j < i; ++j) w[j] = 0;
  }
}

Rule gm_rule(int dim, int degree) {
  FEEC_REQUIRE(dim >= 0 && dim <= 4, FEEC_E_ARG, "quadrature dimension must be 0..4");
  FEEC_REQUIRE(degree >= 0 && degree <= kMaxRuleDegree, FEEC_E_ARG,
               "quadrature degree must be 0..7");
  const int n = dim, s = degree / 2, d = 2 * s + 1;
  auto fact = [](int m) {
    double f = 1.0;
    for (int i = 2; i <= m; ++i) f *= i;
    return f;
  };
  Rule r;
  r.dim = dim;
  std::vector<std::vector<int>> words;
  for (int i = 0; i <= s; ++i) {
    const double den = (double)(d + n - 2 * i);
    double dpow = 1.0;
    for (int e = 0; e < d; ++e) dpow *= den;
    const double wi = ((i & 1) ? -1.0 : 1.0) * std::ldexp(1.0, -2 * s) * dpow / (fact(i) * fact(d + n - i));
    colex_words(n + 1, s - i, words);
    for (const auto& word : words) {
      int parts[5] = {0, 0, 0, 0, 0};
      for (int sym : word) ++parts[sym];
      for (int a = 0; a <= n; ++a) r.bary.push_back((2 * parts[a] + 1) / den);
      r.w.push_back(wi);
    }
  }
  r.np = (int)r.w.size();
  double total = 0.0;
  for (double v : r.w) total += v;
  for (double& v : r.w) v /= total;
  return r;
}

// The rule as one device table: bary [np][dim+1], then the weights.
struct DeviceRule {
  int np = 0;
  DBuf<double> t;
  const double* bary() const { return t.p; }
};
void upload_rule(Ctx* ctx, const Rule& r, DeviceRule& out) {
  std::vector<double> h(r.bary);
  h.insert(h.end(), r.w.begin(), r.w.end());
  out.np = r.np;
  out.t.alloc(ctx, h.size());
  copy_to_device(ctx, out.t.p, h.data(), h.size() * 8);
}

// ---------------------------------------------------------------------------
// per-cell pieces (device)
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void lane_geometry(const int32_t* __restrict__ edges,
                                              const double* __restrict__ len, int64_t c,
                                              CellGeom<D>& cg, double& vol) {
  constexpr int NE = binom(D + 1, 2);
  double s[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) s[e] = len[edges[c * NE + e]];
  cell_geometry<D>(s, cg);
  vol = (1.0 / factorial(D)) * sqrt(fabs(cg.det));
}

// Lambda^K of a D x D matrix on colex axis subsets: out[I][J] = det(a[ax I, ax J]).
template <int D, int K>
__device__ __forceinline__ void lambda_k(const double (&a)[D][D],
                                         double (&out)[Grade<D, K>::NA][Grade<D, K>::NA]) {
  constexpr Grade<D, K> T{};
  constexpr int NA = Grade<D, K>::NA;
#pragma unroll
  for (int I = 0; I < NA; ++I)
#pragma unroll
    for (int J = 0; J < NA; ++J) {
      double m[K > 0 ? K : 1][K > 0 ? K : 1];
#pragma unroll
      for (int x = 0; x < K; ++x)
#pragma unroll
        for (int y = 0; y < K; ++y) m[x][y] = a[T.ax[I][x]][T.ax[J][y]];
      out[I][J] = leibniz<K>(m);
    }
}

// J[a][b] = x_(v_(b+1))[a] - x_(v_0)[a] of cell c (ambient dimension D).
template <int D>
__device__ __forceinline__ void cell_jacobian(const int32_t* __restrict__ verts,
                                              const double* __restrict__ X, int64_t c,
                                              double (&J)[D][D]) {
  const int64_t v0 = verts[c * (D + 1)];
#pragma unroll
  for (int b = 0; b < D; ++b) {
    const int64_t vb = verts[c * (D + 1) + b + 1];
#pragma unroll
    for (int a = 0; a < D; ++a) J[a][b] = X[vb * D + a] - X[v0 * D + a];
  }
}

// The reference-frame value of one node of the field: f itself, or its
// pullback sum_Kamb PB[Kamb][I] f[Kamb] for an ambient-frame field.
template <int D, int K>
__device__ __forceinline__ void node_value(const double* __restrict__ f, bool ambient,
                                           const double (&PB)[Grade<D, K>::NA][Grade<D, K>::NA],
                                           double (&fr)[Grade<D, K>::NA]) {
  constexpr int NA = Grade<D, K>::NA;
  double v[NA];
#pragma unroll
  for (int I = 0; I < NA; ++I) v[I] = f[I];
  if (!ambient) {
#pragma unroll
    for (int I = 0; I < NA; ++I) fr[I] = v[I];
    return;
  }
#pragma unroll
  for (int I = 0; I < NA; ++I) {
    double t = 0.0;
#pragma unroll
    for (int A = 0; A < NA; ++A) t += PB[A][I] * v[A];
    fr[I] = t;
  }
}

// W c at barycentric point lam (reference frame): k! sum_sigma c_sigma
// sum_p (-1)^p lam[sigma_p] P[sigma minus sigma_p].
template <int D, int K>
__device__ __forceinline__ void whitney_value(const double* __restrict__ lam,
                                              const double (&cv)[Grade<D, K>::ND],
                                              double (&u)[Grade<D, K>::NA]) {
  constexpr Grade<D, K> T{};
  constexpr int NA = Grade<D, K>::NA, NP = Grade<D, K>::NP, ND = Grade<D, K>::ND;
  double r[NP];
#pragma unroll
  for (int A = 0; A < NP; ++A) r[A] = 0.0;
#pragma unroll
  for (int i = 0; i < ND; ++i)
#pragma unroll
    for (int p = 0; p <= K; ++p) {
      const double t = cv[i] * lam[T.dv[i][p]];
      r[T.dr[i][p]] += (p & 1) ? -t : t;
    }
#pragma unroll
  for (int I = 0; I < NA; ++I) {
    double t = 0.0;
#pragma unroll
    for (int A = 0; A < NP; ++A)
      if (T.P[A][I] != 0) t += (T.P[A][I] > 0) ? r[A] : -r[A];
    u[I] = factorial(K) * t;
  }
}

// the cochain's values on the cell's grade-K faces, local colex order
template <int D, int K>
__device__ __forceinline__ void cell_cochain(const int32_t* __restrict__ faces,
                                             const double* __restrict__ cochain, int64_t c,
                                             double (&cv)[Grade<D, K>::ND]) {
  constexpr int ND = Grade<D, K>::ND;
#pragma unroll
  for (int i = 0; i < ND; ++i) cv[i] = cochain[K == D ? c : (int64_t)faces[c * ND + i]];
}

// Number of nodes of the Grundmann-Moller rule of index S on the n-simplex.
constexpr int gm_count(int n, int S) {
  int c = 0;
  for (int i = 0; i <= S; ++i) c += binom(S - i + n, n);
  return c;
}
constexpr int kGenericRule = -1;
template <int N, int S>
__host__ __device__ constexpr int rule_nodes(int np) { return S >= 0 ? gm_count(N, S) : np; }
// rows of at most this many doubles are held in registers from the start
template <int S, int W>
constexpr bool kPreload = S >= 0 && W <= 24;

// The block's rows of an output [row][W] array leave LDS as one contiguous
// chunk, consecutive lanes on consecutive doubles (each wavefront store covers
// 512 contiguous bytes). Every thread of the block calls it.
__device__ __forceinline__ void stage_out(double* __restrict__ dst, int64_t row0, int64_t nrows,
                                          int W, const double* lds) {
  __syncthreads();
  const int64_t total = min((int64_t)TPB, nrows - row0) * W;
  double* d = dst + row0 * W;
  for (int64_t i = threadIdx.x; i < total; i += TPB) d[i] = lds[i];
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// Kernels are instantiated for the rule indices S = 0 and 1 (degrees up to 3,
// the reference's defaults), whose node loops unroll completely: a lane issues
// all of its field loads and the rule's (uniform, scalar) loads at once
// instead of waiting on them node by node. S = -1 is the generic instance for
// the higher rules, with the node count np at run time.
//
// Element loads vol * sum_n w_n <f_n, W_sigma(x_n)>_(Lambda^K g^-1), out[sigma][cell].
template <int D, int K, int S>
__global__ void __launch_bounds__(TPB)
    k_source_load(const int32_t* __restrict__ edges, const double* __restrict__ len,
                  const int32_t* __restrict__ verts, const double* __restrict__ X,
                  const double* __restrict__ rule, int np, const double* __restrict__ field,
                  int64_t nc, double* __restrict__ out) {
  constexpr Grade<D, K> T{};
  constexpr int NA = Grade<D, K>::NA, NP = Grade<D, K>::NP, ND = Grade<D, K>::ND;
  constexpr int NQ = S >= 0 ? gm_count(D, S) : 1;
  const int nq = rule_nodes<D, S>(np);
  const int64_t c = blockIdx.x * (int64_t)TPB + threadIdx.x;
  if (c >= nc) return;
  const double* f = field + c * (int64_t)nq * NA;
  // small rows are loaded up front, overlapping the metric's latency
  double fv[kPreload<S, NQ * NA> ? NQ : 1][NA];
  if constexpr (kPreload<S, NQ * NA>) {
#pragma unroll
    for (int n = 0; n < nq; ++n)
#pragma unroll
      for (int I = 0; I < NA; ++I) fv[n][I] = f[n * NA + I];
  }
  CellGeom<D> cg;
  double vol;
  lane_geometry<D>(edges, len, c, cg, vol);
  double L[NA][NA];
  lambda_k<D, K>(cg.ginv, L);
  double PB[NA][NA];
  const bool ambient = X != nullptr;
  if (ambient) {
    double J[D][D];
    cell_jacobian<D>(verts, X, c, J);
    lambda_k<D, K>(J, PB);
  }
  const double* w = rule + nq * (D + 1);
  double acc[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) acc[i] = 0.0;
#pragma unroll
  for (int n = 0; n < nq; ++n) {
    double fr[NA];
    node_value<D, K>(kPreload<S, NQ * NA> ? fv[kPreload<S, NQ * NA> ? n : 0] : f + n * NA, ambient, PB, fr);
    double h[NA];
#pragma unroll
    for (int I = 0; I < NA; ++I) {
      double t = 0.0;
#pragma unroll
      for (int J = 0; J < NA; ++J) t += L[I][J] * fr[J];
      h[I] = t;
    }
    double q[NP];
#pragma unroll
    for (int A = 0; A < NP; ++A) {
      double t = 0.0;
#pragma unroll
      for (int I = 0; I < NA; ++I)
        if (T.P[A][I] != 0) t += (T.P[A][I] > 0) ? h[I] : -h[I];
      q[A] = t;
    }
    const double* lam = rule + n * (D + 1);
    const double wn = w[n];
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      double t = 0.0;
#pragma unroll
      for (int p = 0; p <= K; ++p) {
        const double v = lam[T.dv[i][p]] * q[T.dr[i][p]];
        t += (p & 1) ? -v : v;
      }
      acc[i] += wn * (factorial(K) * t);
    }
  }
#pragma unroll
  for (int i = 0; i < ND; ++i) out[i * nc + c] = vol * acc[i];
}

// W c at every cell node, reference frame: out[cell][node][I], written through
// LDS (dynamic LDS = TPB * NQ * NA doubles when stage).
template <int D, int K, int S>
__global__ void __launch_bounds__(TPB)
    k_whitney_nodes(const int32_t* __restrict__ faces, const double* __restrict__ cochain,
                    const double* __restrict__ rule, int np, int64_t nc, bool stage,
                    double* __restrict__ out) {
  constexpr int NA = Grade<D, K>::NA, ND = Grade<D, K>::ND;
  const int nq = rule_nodes<D, S>(np);
  extern __shared__ double lds[];
  const int64_t c0 = blockIdx.x * (int64_t)TPB;
  const int64_t c = c0 + threadIdx.x;
  double* o = stage ? lds + threadIdx.x * nq * NA : out + c * (int64_t)nq * NA;
  if (c < nc) {
    double cv[ND];
    cell_cochain<D, K>(faces, cochain, c, cv);
#pragma unroll
    for (int n = 0; n < nq; ++n) {
      double u[NA];
      whitney_value<D, K>(rule + n * (D + 1), cv, u);
#pragma unroll
      for (int I = 0; I < NA; ++I) o[n * NA + I] = u[I];
    }
  }
  if (stage) stage_out(out, c0, nc, nq * NA, lds);
}

// Per cell vol * sum_n w_n <e_n, e_n>_(Lambda^K g^-1), e = f - W c, summed
// per block (fixed order) into part[blockIdx.x].
template <int D, int K, int S>
__global__ void __launch_bounds__(TPB)
    k_l2_error(const int32_t* __restrict__ edges, const double* __restrict__ len,
               const int32_t* __restrict__ verts, const double* __restrict__ X,
               const int32_t* __restrict__ faces, const double* __restrict__ cochain,
               const double* __restrict__ rule, int np, const double* __restrict__ field,
               int64_t nc, double* __restrict__ part) {
  constexpr int NA = Grade<D, K>::NA, ND = Grade<D, K>::ND;
  constexpr int NQ = S >= 0 ? gm_count(D, S) : 1;
  const int nq = rule_nodes<D, S>(np);
  __shared__ double red[TPB];
  const int64_t c = blockIdx.x * (int64_t)TPB + threadIdx.x;
  double cell = 0.0;
  if (c < nc) {
    const double* f = field + c * (int64_t)nq * NA;
    double fv[kPreload<S, NQ * NA> ? NQ : 1][NA];
    if constexpr (kPreload<S, NQ * NA>) {
#pragma unroll
      for (int n = 0; n < nq; ++n)
#pragma unroll
        for (int I = 0; I < NA; ++I) fv[n][I] = f[n * NA + I];
    }
    double cv[ND];
    cell_cochain<D, K>(faces, cochain, c, cv);
    CellGeom<D> cg;
    double vol;
    lane_geometry<D>(edges, len, c, cg, vol);
    double L[NA][NA];
    lambda_k<D, K>(cg.ginv, L);
    double PB[NA][NA];
    const bool ambient = X != nullptr;
    if (ambient) {
      double J[D][D];
      cell_jacobian<D>(verts, X, c, J);
      lambda_k<D, K>(J, PB);
    }
    const double* w = rule + nq * (D + 1);
    double acc = 0.0;
#pragma unroll
    for (int n = 0; n < nq; ++n) {
      double fr[NA], u[NA];
      node_value<D, K>(kPreload<S, NQ * NA> ? fv[kPreload<S, NQ * NA> ? n : 0] : f + n * NA, ambient, PB,
                       fr);
      whitney_value<D, K>(rule + n * (D + 1), cv, u);
      double e[NA];
#pragma unroll
      for (int I = 0; I < NA; ++I) e[I] = fr[I] - u[I];
      double q = 0.0;
#pragma unroll
      for (int I = 0; I < NA; ++I) {
        double t = 0.0;
#pragma unroll
        for (int J = 0; J < NA; ++J) t += L[I][J] * e[J];
        q += e[I] * t;
      }
      acc += w[n] * q;
    }
    cell = vol * acc;
  }
  red[threadIdx.x] = cell;
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Vertices of grade-K simplex s (ascending): the vertex itself, a cell's
// vertices, or decoded from the representative slot cell * nf + f.
template <int D, int K>
__device__ __forceinline__ void simplex_vertices(const int32_t* __restrict__ verts,
                                                 const int64_t* __restrict__ rep, int64_t s,
                                                 int64_t (&v)[K + 1]) {
  if constexpr (K == 0) {
    v[0] = s;
  } else if constexpr (K == D) {
#pragma unroll
    for (int p = 0; p <= K; ++p) v[p] = verts[s * (D + 1) + p];
  } else {
    constexpr int NF = binom(D + 1, K + 1);
    const int64_t slot = rep[s];
    const int64_t cell = slot / NF;
    const unsigned m = nth_subset(D + 1, K + 1, (int)(slot % NF));
#pragma unroll
    for (int p = 0; p <= K; ++p) v[p] = verts[cell * (D + 1) + nth_elem(m, p)];
  }
}

// Node positions sum_i lam_i x_(v_i) of every grade-K simplex (K = D: cells),
// written through LDS like k_whitney_nodes.
template <int D, int K, int S>
__global__ void __launch_bounds__(TPB)
    k_simplex_points(const int32_t* __restrict__ verts, const int64_t* __restrict__ rep,
                     const double* __restrict__ X, const double* __restrict__ rule, int np,
                     int64_t ns, bool stage, double* __restrict__ out) {
  const int nq = rule_nodes<K, S>(np);
  extern __shared__ double lds[];
  const int64_t s0 = blockIdx.x * (int64_t)TPB;
  const int64_t s = s0 + threadIdx.x;
  double* o = stage ? lds + threadIdx.x * nq * D : out + s * (int64_t)nq * D;
  if (s < ns) {
    int64_t v[K + 1];
    simplex_vertices<D, K>(verts, rep, s, v);
    double x[K + 1][D];
#pragma unroll
    for (int i = 0; i <= K; ++i)
#pragma unroll
      for (int a = 0; a < D; ++a) x[i][a] = X[v[i] * D + a];
#pragma unroll
    for (int n = 0; n < nq; ++n) {
      const double* lam = rule + n * (K + 1);
#pragma unroll
      for (int a = 0; a < D; ++a) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i <= K; ++i) t += lam[i] * x[i][a];
        o[n * D + a] = t;
      }
    }
  }
  if (stage) stage_out(out, s0, ns, nq * D, lds);
}

// c_s = (1/K!) sum_n w_n sum_Kamb f[s][n][Kamb] det(E[ax Kamb, :]), E the
// simplex's edge vectors x_(s_(b+1)) - x_(s_0) (D x K).
template <int D, int K, int S>
__global__ void __launch_bounds__(TPB)
    k_derham(const int32_t* __restrict__ verts, const int64_t* __restrict__ rep,
             const double* __restrict__ X, const double* __restrict__ rule, int np,
             const double* __restrict__ field, int64_t ns, double* __restrict__ out) {
  constexpr Grade<D, K> T{};
  constexpr int NA = Grade<D, K>::NA;
  const int nq = rule_nodes<K, S>(np);
  const int64_t s = blockIdx.x * (int64_t)TPB + threadIdx.x;
  if (s >= ns) return;
  const double* f = field + s * (int64_t)nq * NA;
  int64_t v[K + 1];
  simplex_vertices<D, K>(verts, rep, s, v);
  double blade[NA];
#pragma unroll
  for (int A = 0; A < NA; ++A) {
    double m[K > 0 ? K : 1][K > 0 ? K : 1];
#pragma unroll
    for (int x = 0; x < K; ++x)
#pragma unroll
      for (int b = 0; b < K; ++b) {
        const int a = T.ax[A][x];
        m[x][b] = X[v[b + 1] * D + a] - X[v[0] * D + a];
      }
    blade[A] = leibniz<K>(m);
  }
  const double* w = rule + nq * (K + 1);
  double acc = 0.0;
#pragma unroll
  for (int n = 0; n < nq; ++n) {
    double t = 0.0;
#pragma unroll
    for (int A = 0; A < NA; ++A) t += f[n * NA + A] * blade[A];
    acc += w[n] * t;
  }
  out[s] = (1.0 / factorial(K)) * acc;
}

__global__ void k_zero_i32(int32_t* p, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_load_srcs(int64_t nc, int nf, uint32_t* srcs) {
  const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
  if (i < nc * nf) srcs[i] = (uint32_t)((i % nf) * nc + i / nf);
}

__global__ void k_inv_abs_diag(const double* d, int64_t n, bool absolute, double* out) {
  const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
  if (i < n) out[i] = d[i] != 0.0 ? 1.0 / (absolute ? fabs(d[i]) : d[i]) : 1.0;
}

// ---------------------------------------------------------------------------
// dispatch over (D, K) and the rule index S
// ---------------------------------------------------------------------------
template <template <int, int, int> class F, int D, int K, class... A>
void by_s(int S, A&&... a) {
  switch (S) {
    case 0: F<D, K, 0>::run(a...); return;
    case 1: F<D, K, 1>::run(a...); return;
    default: F<D, K, kGenericRule>::run(a...); return;
  }
}

template <template <int, int, int> class F, class... A>
void by_dks(int d, int k, int S, A&&... a) {
  switch (d * 8 + k) {
#define FEEC_DK(D, K)                \
  case D * 8 + K:                    \
    by_s<F, D, K>(S, a...);          \
    return;
    FEEC_DK(1, 0) FEEC_DK(1, 1)
    FEEC_DK(2, 0) FEEC_DK(2, 1) FEEC_DK(2, 2)
    FEEC_DK(3, 0) FEEC_DK(3, 1) FEEC_DK(3, 2) FEEC_DK(3, 3)
    FEEC_DK(4, 0) FEEC_DK(4, 1) FEEC_DK(4, 2) FEEC_DK(4, 3) FEEC_DK(4, 4)
#undef FEEC_DK
  }
  fail(FEEC_E_UNSUPPORTED, "dimension 1..4 and grade 0..dim");
}

// the dynamic LDS of a block that writes rows of W doubles through LDS, 0 when
// that chunk would exceed kStageBytes (then the rows are written directly)
constexpr int kStageBytes = 48 * 1024;
inline size_t stage_bytes(int W) {
  const size_t b = (size_t)TPB * W * 8;
  return b <= (size_t)kStageBytes ? b : 0;
}

template <int D, int K, int S>
struct RunLoad {
  static void run(Ctx* ctx, const int32_t* edges, const double* len, const int32_t* verts,
                  const double* X, const DeviceRule& r, const double* field, int64_t nc,
                  double* out) {
    FEEC_REQUIRE((r.np == rule_nodes<D, S>(r.np)), FEEC_E_ARG, "rule does not match its index");
    k_source_load<D, K, S><<<grid_for(nc), TPB, 0, ctx->stream>>>(edges, len, verts, X, r.t.p,
                                                                   r.np, field, nc, out);
    FEEC_HIP(hipGetLastError());
  }
};
template <int D, int K, int S>
struct RunWhitney {
  static void run(Ctx* ctx, const int32_t* faces, const double* cochain, const DeviceRule& r,
                  int64_t nc, double* out) {
    FEEC_REQUIRE((r.np == rule_nodes<D, S>(r.np)), FEEC_E_ARG, "rule does not match its index");
    const size_t lds = stage_bytes(r.np * Grade<D, K>::NA);
    k_whitney_nodes<D, K, S><<<grid_for(nc), TPB, lds, ctx->stream>>>(faces, cochain, r.t.p, r.np,
                                                                       nc, lds != 0, out);
    FEEC_HIP(hipGetLastError());
  }
};
template <int D, int K, int S>
struct RunError {
  static void run(Ctx* ctx, const int32_t* edges, const double* len, const int32_t* verts,
                  const double* X, const int32_t* faces, const double* cochain,
                  const DeviceRule& r, const double* field, int64_t nc, double* part) {
    FEEC_REQUIRE((r.np == rule_nodes<D, S>(r.np)), FEEC_E_ARG, "rule does not match its index");
    k_l2_error<D, K, S><<<grid_for(nc), TPB, 0, ctx->stream>>>(edges, len, verts, X, faces,
                                                                cochain, r.t.p, r.np, field, nc,
                                                                part);
    FEEC_HIP(hipGetLastError());
  }
};
template <int D, int K, int S>
struct RunPoints {
  static void run(Ctx* ctx, const int32_t* verts, const int64_t* rep, const double* X,
                  const DeviceRule& r, int64_t ns, double* out) {
    FEEC_REQUIRE((r.np == rule_nodes<K, S>(r.np)), FEEC_E_ARG, "rule does not match its index");
    const size_t lds = stage_bytes(r.np * D);
    k_simplex_points<D, K, S><<<grid_for(ns), TPB, lds, ctx->stream>>>(verts, rep, X, r.t.p, r.np,
                                                                        ns, lds != 0, out);
    FEEC_HIP(hipGetLastError());
  }
};
template <int D, int K, int S>
struct RunDerham {
  static void run(Ctx* ctx, const int32_t* verts, const int64_t* rep, const double* X,
                  const DeviceRule& r, const double* field, int64_t ns, double* out) {
    FEEC_REQUIRE((r.np == rule_nodes<K, S>(r.np)), FEEC_E_ARG, "rule does not match its index");
    k_derham<D, K, S><<<grid_for(ns), TPB, 0, ctx->stream>>>(verts, rep, X, r.t.p, r.np, field, ns,
                                                              out);
    FEEC_HIP(hipGetLastError());
  }
};

// ---------------------------------------------------------------------------
// host plumbing
// ---------------------------------------------------------------------------
int64_t ncomp(int d, int k) { return binom(d, k); }

}  // namespace

// the geometry's coordinates on the device, uploaded on first use
const double* device_coords(Geom& g) {
  Mesh& m = *g.mesh;
  FEEC_REQUIRE(!g.coords.empty(), FEEC_E_ARG,
               "an ambient frame needs a geometry made from coordinates");
  std::lock_guard<std::mutex> lock(m.mu);
  if (!g.dcoords.p) {
    g.dcoords.alloc(m.ctx, g.coords.size());
    copy_to_device(m.ctx, g.dcoords.p, g.coords.data(), g.coords.size() * 8);
  }
  return g.dcoords.p;
}

namespace {

// An input array: the caller's device pointer, or a device copy of its host array.
struct In {
  const double* p = nullptr;
  DBuf<double> own;
  In(Ctx* ctx, const double* src, int64_t n, bool on_device) {
    if (on_device || n == 0) {
      p = src;
      return;
    }
    own.alloc(ctx, n);
    copy_to_device(ctx, own.p, src, n * 8);
    p = own.p;
  }
};
// An output array: the caller's device pointer, or a device buffer copied back by done().
struct Out {
  double* p = nullptr;
  double* host = nullptr;
  int64_t n = 0;
  DBuf<double> own;
  Out(Ctx* ctx, double* dst, int64_t count, bool on_device) : n(count) {
    if (on_device) {
      p = dst;
      return;
    }
    host = dst;
    own.alloc(ctx, std::max<int64_t>(count, 1));
    p = own.p;
  }
  void done(Ctx* ctx) {
    if (host && n) copy_to_host(ctx, host, own.p, n * 8);
    else ctx->sync();
  }
};

void check_grade(const Mesh& m, int grade) {
  FEEC_REQUIRE(m.d >= 1 && m.d <= 4, FEEC_E_UNSUPPORTED, "dimension 1..4");
  FEEC_REQUIRE(grade >= 0 && grade <= m.d, FEEC_E_ARG, "grade out of range");
}
bool frame_ambient(const feec_frame& frame) {
  const int v = abi_int(frame);
  FEEC_REQUIRE(v == FEEC_FRAME_REFERENCE || v == FEEC_FRAME_AMBIENT, FEEC_E_ARG,
               "frame must be a feec_frame value");
  return v == FEEC_FRAME_AMBIENT;
}

// The cached plan that sums element loads [slot][cell] per grade-k face, in
// cell order (k < d).
const CooPlan& load_plan(Mesh& m, int k) {
  std::lock_guard<std::mutex> lock(m.mu);
  auto key = std::make_tuple(kLoadPlanTag, k, 0);
  auto it = m.plans.find(key);
  if (it != m.plans.end()) return *it->second;
  Ctx* ctx = m.ctx;
  const int64_t nc = m.n[m.d];
  const int nf = m.nfaces(k);
  const int64_t total = nc * nf;
  FEEC_REQUIRE(total < (int64_t(1) << 31), FEEC_E_ARG, "too many cell faces for one load");
  DBuf<int32_t> cols(ctx, total);
  DBuf<uint32_t> srcs(ctx, total);
  k_zero_i32<<<grid_for(total), TPB, 0, ctx->stream>>>(cols.p, total);
  k_load_srcs<<<grid_for(total), TPB, 0, ctx->stream>>>(nc, nf, srcs.p);
  FEEC_HIP(hipGetLastError());
  auto plan = std::make_unique<CooPlan>();
  coo_plan_build(ctx, *plan, m.n[k], 1, total, m.faces[k].p, cols.p, srcs.p);
  FEEC_REQUIRE(plan->nnz == m.n[k], FEEC_E_TOPOLOGY, "a face in no cell");
  const CooPlan& ref = *plan;
  m.plans[key] = std::move(plan);
  return ref;
}

// The load vector of a field given at the cell nodes of `degree` (device in,
// device out, N_k entries).
void source_load_dev(Geom& g, int k, int degree, bool ambient, const double* field, double* out) {
  Mesh& m = *g.mesh;
  Ctx* ctx = m.ctx;
  const int d = m.d;
  const int64_t nc = m.n[d];
  Rule r = gm_rule(d, degree);
  DeviceRule dr;
  upload_rule(ctx, r, dr);
  const double* X = ambient ? device_coords(g) : nullptr;
  const int nf = m.nfaces(k);
  if (k == d) {
    by_dks<RunLoad>(d, k, degree / 2, ctx, m.faces[1].p, g.lengths.p, m.faces[0].p, X, dr, field, nc, out);
    ctx->sync();
    return;
  }
  DBuf<double> el(ctx, std::max<int64_t>(nc * nf, 1));
  by_dks<RunLoad>(d, k, degree / 2, ctx, m.faces[1].p, g.lengths.p, m.faces[0].p, X, dr, field, nc, el.p);
  const CooPlan& plan = load_plan(m, k);
  coo_plan_values(ctx, plan, el.p, out, nullptr);
  ctx->sync();
}

// A device copy of a CSR (same context).
void csr_clone(const Csr& a, Csr& out) {
  Ctx* ctx = a.ctx;
  out.ctx = ctx;
  out.nrows = a.nrows;
  out.ncols = a.ncols;
  out.nnz = a.nnz;
  out.row_ptr.alloc(ctx, a.nrows + 1);
  out.col.alloc(ctx, a.nnz);
  out.val.alloc(ctx, a.nnz);
  FEEC_HIP(hipMemcpyAsync(out.row_ptr.p, a.row_ptr.p, (a.nrows + 1) * 8, hipMemcpyDeviceToDevice,
                          ctx->stream));
  if (a.nnz) {
    FEEC_HIP(hipMemcpyAsync(out.col.p, a.col.p, a.nnz * 4, hipMemcpyDeviceToDevice, ctx->stream));
    FEEC_HIP(hipMemcpyAsync(out.val.p, a.val.p, a.nnz * 8, hipMemcpyDeviceToDevice, ctx->stream));
  }
}

}  // namespace

}  // namespace feec

using namespace feec;

extern "C" {

feec_status feec_quadrature_rule(int dim, int degree, int* npoints, double* bary,
                                 double* weights) {
  return guard([&] {
    FEEC_REQUIRE(npoints, FEEC_E_ARG, "null argument");
    Rule r = gm_rule(dim, degree);
    *npoints = r.np;
    if (bary) std::copy(r.bary.begin(), r.bary.end(), bary);
    if (weights) std::copy(r.w.begin(), r.w.end(), weights);
  });
}

feec_status feec_cell_quadrature_points(feec_geom* g, int degree, double* points, int on_device) {
  return guard([&] {
    FEEC_REQUIRE(g && points, FEEC_E_ARG, "null argument");
    Mesh& m = *g->mesh;
    Ctx* ctx = m.ctx;
    FEEC_HIP(hipSetDevice(ctx->device));
    check_grade(m, m.d);
    Rule r = gm_rule(m.d, degree);
    DeviceRule dr;
    upload_rule(ctx, r, dr);
    const double* X = device_coords(*g);
    const int64_t nc = m.n[m.d];
    Out o(ctx, points, nc * r.np * m.d, on_device != 0);
    by_dks<RunPoints>(m.d, m.d, degree / 2, ctx, m.faces[0].p, (const int64_t*)nullptr, X, dr, nc, o.p);
    o.done(ctx);
  });
}

feec_status feec_simplex_quadrature_points(feec_geom* g, int grade, int degree, double* points,
                                           int on_device) {
  return guard([&] {
    FEEC_REQUIRE(g && points, FEEC_E_ARG, "null argument");
    Mesh& m = *g->mesh;
    Ctx* ctx = m.ctx;
    FEEC_HIP(hipSetDevice(ctx->device));
    check_grade(m, grade);
    Rule r = gm_rule(grade, degree);
    DeviceRule dr;
    upload_rule(ctx, r, dr);
    const double* X = device_coords(*g);
    const int64_t ns = m.n[grade];
    Out o(ctx, points, ns * r.np * m.d, on_device != 0);
    const int64_t* rep = (grade > 0 && grade < m.d) ? m.rep[grade].p : nullptr;
    by_dks<RunPoints>(m.d, grade, degree / 2, ctx, m.faces[0].p, rep, X, dr, ns, o.p);
    o.done(ctx);
  });
}

feec_status feec_source_load(feec_geom* g, int grade, int degree, feec_frame frame,
                             const double* field, double* load, int on_device) {
  return guard([&] {
    FEEC_REQUIRE(g && field && load, FEEC_E_ARG, "null argument");
    const bool ambient = frame_ambient(frame);
    Mesh& m = *g->mesh;
    Ctx* ctx = m.ctx;
    FEEC_HIP(hipSetDevice(ctx->device));
    check_grade(m, grade);
    const int np = gm_rule(m.d, degree).np;
    In f(ctx, field, m.n[m.d] * np * ncomp(m.d, grade), on_device != 0);
    Out o(ctx, load, m.n[grade], on_device != 0);
    source_load_dev(*g, grade, degree, ambient, f.p, o.p);
    o.done(ctx);
  });
}

feec_status feec_l2_projection(feec_geom* g, int grade, int degree, feec_frame frame,
                               const double* field, const feec_csr* mass, double* cochain,
                               feec_report* report, int on_device) {
  return guard([&] {
    FEEC_REQUIRE(g && field && cochain, FEEC_E_ARG, "null argument");
    const bool ambient = frame_ambient(frame);
    Mesh& m = *g->mesh;
    Ctx* ctx = m.ctx;
    FEEC_HIP(hipSetDevice(ctx->device));
    check_grade(m, grade);
    const int np = gm_rule(m.d, degree).np;
    const int64_t n = m.n[grade];
    In f(ctx, field, m.n[m.d] * np * ncomp(m.d, grade), on_device != 0);
    DBuf<double> b(ctx, std::max<int64_t>(n, 1));
    source_load_dev(*g, grade, degree, ambient, f.p, b.p);
    Csr own;
    if (mass) {
      FEEC_REQUIRE(mass->ctx == ctx && mass->nrows == n && mass->ncols == n, FEEC_E_ARG,
                   "mass must be this geometry's N_k x N_k matrix");
    } else {
      assemble_form(*g, FEEC_FULL, FEEC_MASS, grade, own);
    }
    const Csr& M = mass ? *static_cast<const Csr*>(mass) : own;
    // Riemannian on every cell: the mass is SPD and Jacobi CG is the reference's solve
    bool riemannian = g->euclidean;
    if (!riemannian) {
      int64_t counts[3];
      DBuf<uint8_t> flags(ctx, std::max<int64_t>(m.n[m.d], 1));
      causal_census(*g, counts, flags.p);
      std::vector<uint8_t> h(m.n[m.d]);
      copy_to_host(ctx, h.data(), flags.p, h.size());
      riemannian = std::all_of(h.begin(), h.end(), [](uint8_t v) { return v != 0; });
    }
    Out o(ctx, cochain, n, on_device != 0);
    DBuf<double> diag(ctx, std::max<int64_t>(n, 1)), dinv(ctx, std::max<int64_t>(n, 1));
    csr_diagonal(M, diag.p);
    k_inv_abs_diag<<<grid_for(n), TPB, 0, ctx->stream>>>(diag.p, n, !riemannian, dinv.p);
    FEEC_HIP(hipGetLastError());
    feec_report rep{};
    if (riemannian) {
      KrylovStats ks = cg_solve_device(ctx, M, dinv.p, b.p, o.p, kProjectionRtol, kProjectionMaxIters);
      rep.iters = ks.iters;
      rep.converged = ks.converged ? 1 : 0;
      rep.residual = ks.precond_residual;
      rep.precond_residual = ks.precond_residual;
    } else {
      // symmetric indefinite mass: MINRES with the inverse absolute diagonal
      KktSystem sys;
      sys.ctx = ctx;
      sys.n = n;
      sys.nh = 0;
      if (mass) {
        csr_clone(M, sys.A);
      } else {
        sys.A = std::move(own);
      }
      sys.inv_diag = std::move(dinv);
      feec_solve_opts opts;
      feec_solve_opts_default(&opts);
      opts.rtol = kProjectionRtol;
      opts.precision = FEEC_PREC_FP64;
      opts.max_iters = kProjectionMaxIters;
      SolverState* ws = nullptr;
      rep = kkt_solve(sys, b.p, o.p, opts, &ws);
      solver_state_free(ws);
    }
    o.done(ctx);
    if (report) *report = rep;
    FEEC_REQUIRE(rep.converged, FEEC_E_NOT_CONVERGED, "the mass solve did not reach 1e-12");
  });
}

feec_status feec_whitney_at_nodes(feec_geom* g, int grade, int degree, const double* cochain,
                                  double* values, int on_device) {
  return guard([&] {
    FEEC_REQUIRE(g && cochain && values, FEEC_E_ARG, "null argument");
    Mesh& m = *g->mesh;
    Ctx* ctx = m.ctx;
    FEEC_HIP(hipSetDevice(ctx->device));
    check_grade(m, grade);
    Rule r = gm_rule(m.d, degree);
    DeviceRule dr;
    upload_rule(ctx, r, dr);
    const int64_t nc = m.n[m.d];
    In c(ctx, cochain, m.n[grade], on_device != 0);
    Out o(ctx, values, nc * r.np * ncomp(m.d, grade), on_device != 0);
    const int32_t* faces = grade < m.d ? m.faces[grade].p : nullptr;
    by_dks<RunWhitney>(m.d, grade, degree / 2, ctx, faces, c.p, dr, nc, o.p);
    o.done(ctx);
  });
}

feec_status feec_l2_error(feec_geom* g, int grade, const double* cochain, feec_frame frame,
                          const double* exact, double* error, int on_device) {
  return guard([&] {
    FEEC_REQUIRE(g && cochain && exact && error, FEEC_E_ARG, "null argument");
    const bool ambient = frame_ambient(frame);
    Mesh& m = *g->mesh;
    Ctx* ctx = m.ctx;
    FEEC_HIP(hipSetDevice(ctx->device));
    check_grade(m, grade);
    // fe_l2_error integrates with the degree-3 rule (fe.rs:57)
    Rule r = gm_rule(m.d, 3);
    DeviceRule dr;
    upload_rule(ctx, r, dr);
    const int64_t nc = m.n[m.d];
    In c(ctx, cochain, m.n[grade], on_device != 0);
    In f(ctx, exact, nc * r.np * ncomp(m.d, grade), on_device != 0);
    const double* X = ambient ? device_coords(*g) : nullptr;
    const unsigned nb = grid_for(nc);
    DBuf<double> part(ctx, nb);
    const int32_t* faces = grade < m.d ? m.faces[grade].p : nullptr;
    by_dks<RunError>(m.d, grade, 1, ctx, m.faces[1].p, g->lengths.p, m.faces[0].p, X, faces, c.p, dr,
                    f.p, nc, part.p);
    std::vector<double> h(nb);
    copy_to_host(ctx, h.data(), part.p, nb * 8);
    double total = 0.0;
    for (double v : h) total += v;
    *error = std::sqrt(total);
  });
}

feec_status feec_derham_map(feec_geom* g, int grade, int degree, const double* field,
                            double* cochain, int on_device) {
  return guard([&] {
    FEEC_REQUIRE(g && field && cochain, FEEC_E_ARG, "null argument");
    Mesh& m = *g->mesh;
    Ctx* ctx = m.ctx;
    FEEC_HIP(hipSetDevice(ctx->device));
    check_grade(m, grade);
    Rule r = gm_rule(grade, degree);
    DeviceRule dr;
    upload_rule(ctx, r, dr);
    const double* X = device_coords(*g);
    const int64_t ns = m.n[grade];
    In f(ctx, field, ns * r.np * ncomp(m.d, grade), on_device != 0);
    Out o(ctx, cochain, ns, on_device != 0);
    const int64_t* rep = (grade > 0 && grade < m.d) ? m.rep[grade].p : nullptr;
    by_dks<RunDerham>(m.d, grade, degree / 2, ctx, m.faces[0].p, rep, X, dr, f.p, ns, o.p);
    o.done(ctx);
  });
}

}  // extern "C"
