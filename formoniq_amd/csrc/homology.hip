// Exact (co)homology of the simplicial complex, and the harmonic bases built on
// it for meshes that are not Kuhn boxes.
//
// Reference behaviour (luiswirth/formoniq, crates/):
//   IntegerMatrix::rank, rank_mod            simplicial/src/linalg/exact.rs:97-147
//   kernel (RREF, one vector per free column) exact.rs:149-175, 266-298
//   quotient_generators, EchelonSpan         exact.rs:178-264
//   primitive                                exact.rs:300-330
//   Complex::betti_number                    simplicial/src/topology/homology.rs:47-52
//   relative_betti_number                    homology.rs:81-98
//   (relative_)homology_generators           homology.rs:116-153
//   (relative_)cohomology_generators         simplicial/src/topology/cohomology.rs:55-90
//   pairing                                  simplicial/src/topology/chain.rs:358-375
//   harmonics, period_normalize              formoniq/src/harmonic.rs:91-175
//   solve_harmonics                          formoniq/src/problems/elliptic.rs:197-220
//
// Everything integer here runs on the host: it is sparse elimination over a
// prime field or over Q, sequential by nature and sized by the mesh's
// topology, not by its DOFs' arithmetic. The floating-point half of the
// harmonic basis (the mass, the normal equations and their CG) runs on the
// device.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <future>
#include <vector>

#include "feec_internal.hpp"

namespace feec {

namespace {

// A bound on the entries the eliminations may touch (merge steps), so that a
// mesh whose elimination would fill in without limit fails in bounded time.
constexpr int64_t kExactWork = int64_t(4) << 30;

struct Work {
  int64_t used = 0;
  void add(int64_t n) {
    used += n;
    if (used > kExactWork)
      fail(FEEC_E_UNSUPPORTED,
           "exact (co)homology of this mesh exceeds the elimination budget (meshes other than "
           "Kuhn boxes are limited to what sparse elimination finishes quickly)");
  }
};

// ---- exact rationals with 64-bit numerators and denominators ---------------
[[noreturn]] void overflow() {
  fail(FEEC_E_UNSUPPORTED, "an exact (co)homology coefficient exceeds 64 bits");
}

__int128 abs128(__int128 a) { return a < 0 ? -a : a; }

__int128 gcd128(__int128 a, __int128 b) {
  a = abs128(a);
  b = abs128(b);
  while (b) {
    const __int128 t = a % b;
    a = b;
    b = t;
  }
  return a;
}

struct Rat {
  int64_t n = 0, d = 1;
  bool zero() const { return n == 0; }
};

Rat rat(__int128 n, __int128 d) {
  if (d < 0) {
    n = -n;
    d = -d;
  }
  const __int128 g = gcd128(n, d);
  if (g > 1) {
    n /= g;
    d /= g;
  }
  if (n > INT64_MAX || n < -INT64_MAX || d > INT64_MAX) overflow();
  return Rat{(int64_t)n, (int64_t)d};
}
Rat operator-(Rat a, Rat b) { return rat((__int128)a.n * b.d - (__int128)b.n * a.d, (__int128)a.d * b.d); }
Rat operator+(Rat a, Rat b) { return rat((__int128)a.n * b.d + (__int128)b.n * a.d, (__int128)a.d * b.d); }
Rat operator*(Rat a, Rat b) { return rat((__int128)a.n * b.n, (__int128)a.d * b.d); }
Rat operator/(Rat a, Rat b) { return rat((__int128)a.n * b.d, (__int128)a.d * b.n); }

// ---- sparse rows ------------------------------------------------------------
template <class T>
using Row = std::vector<std::pair<int32_t, T>>;  // sorted by position, no zeros

bool is_zero(int64_t v) { return v == 0; }
bool is_zero(const Rat& v) { return v.zero(); }

// out = a - c * b (rows sorted by position), with `sub(x, y)` = x - c y and
// `neg(y)` = -c y on the values
template <class T, class Sub, class Neg>
void row_axpy(const Row<T>& a, const Row<T>& b, Row<T>& out, Sub sub, Neg neg) {
  out.clear();
  size_t i = 0, j = 0;
  while (i < a.size() || j < b.size()) {
    if (j == b.size() || (i < a.size() && a[i].first < b[j].first)) {
      out.push_back(a[i++]);
    } else if (i == a.size() || b[j].first < a[i].first) {
      T v = neg(b[j].second);
      if (!is_zero(v)) out.push_back({b[j].first, v});
      ++j;
    } else {
      T v = sub(a[i].second, b[j].second);
      if (!is_zero(v)) out.push_back({a[i].first, v});
      ++i;
      ++j;
    }
  }
}

int64_t mod_inverse(int64_t a, int64_t p) {
  int64_t t = 0, nt = 1, r = p, nr = a % p;
  while (nr) {
    const int64_t q = r / nr;
    int64_t x = t - q * nt;
    t = nt;
    nt = x;
    x = r - q * nr;
    r = nr;
    nr = x;
  }
  return t < 0 ? t + p : t;
}

// Rank over F_p (exact.rs:117-147): each row reduced against the pivot rows
// keyed by their leading position until it vanishes or opens a new pivot.
int64_t rank_mod(const IntMat& a, int64_t p, Work& w) {
  std::vector<int32_t> pivot_of(std::max<int64_t>(a.ncols, 1), -1);
  std::vector<Row<int64_t>> piv;
  Row<int64_t> row, tmp;
  for (const auto& r0 : a.rows) {
    row.clear();
    for (const auto& e : r0) {
      const int64_t x = ((e.second % p) + p) % p;
      if (x) row.push_back({e.first, x});
    }
    while (!row.empty()) {
      const int32_t lead = row[0].first;
      const int64_t c = row[0].second;
      if (pivot_of[lead] < 0) {
        const int64_t inv = mod_inverse(c, p);
        for (auto& e : row) e.second = e.second * inv % p;
        pivot_of[lead] = (int32_t)piv.size();
        piv.push_back(row);
        break;
      }
      const Row<int64_t>& pv = piv[pivot_of[lead]];
      w.add((int64_t)(row.size() + pv.size()));
      row_axpy<int64_t>(
          row, pv, tmp, [&](int64_t x, int64_t y) { return ((x - c * y % p) % p + p) % p; },
          [&](int64_t y) { return (p - c * y % p) % p; });
      row.swap(tmp);
    }
  }
  return (int64_t)piv.size();
}

// the reduced integer vector along a rational one: denominators cleared, the
// gcd divided out, the first nonzero entry positive (exact.rs:300-330)
std::vector<int64_t> primitive(const std::vector<Rat>& v) {
  __int128 l = 1;
  for (const Rat& x : v)
    if (!x.zero()) {
      l = l / gcd128(l, x.d) * x.d;
      if (l > INT64_MAX) overflow();
    }
  std::vector<__int128> ints(v.size());
  __int128 g = 0;
  for (size_t i = 0; i < v.size(); ++i) {
    ints[i] = (__int128)v[i].n * (l / v[i].d);
    g = gcd128(g, ints[i]);
  }
  std::vector<int64_t> out(v.size(), 0);
  int sign = 0;
  for (size_t i = 0; i < v.size(); ++i) {
    __int128 x = g ? ints[i] / g : ints[i];
    if (!sign && x) sign = x < 0 ? -1 : 1;
    ints[i] = x;
  }
  for (size_t i = 0; i < v.size(); ++i) {
    const __int128 x = sign < 0 ? -ints[i] : ints[i];
    if (x > INT64_MAX || x < -INT64_MAX) overflow();
    out[i] = (int64_t)x;
  }
  return out;
}

// The integer coboundary D_k: C^k -> C^{k+1} on the complex's DOFs (relative:
// the interior simplices, renumbered), rows the (k+1)-DOFs; empty (no rows or
// no columns) outside 0 <= k < d.
IntMat coboundary_int(Mesh& m, int bc, int k) {
  const bool rel = bc == FEEC_RELATIVE;
  if (rel) m.ensure_interior();
  auto dofs = [&](int j) -> int64_t {
    if (j < 0 || j > m.d) return 0;
    return rel ? m.n_interior[j] : m.n[j];
  };
  IntMat a;
  a.nrows = dofs(k + 1);
  a.ncols = dofs(k);
  a.rows.assign(a.nrows, {});
  if (k < 0 || k >= m.d || !a.nrows || !a.ncols) return a;
  Csr full;
  mesh_coboundary(m, k, full);
  std::vector<int64_t> rp(full.nrows + 1);
  std::vector<int32_t> col(full.nnz);
  std::vector<double> val(full.nnz);
  copy_to_host(m.ctx, rp.data(), full.row_ptr.p, rp.size() * 8);
  copy_to_host(m.ctx, col.data(), full.col.p, col.size() * 4);
  copy_to_host(m.ctx, val.data(), full.val.p, val.size() * 8);
  std::vector<int32_t> rmap, cmap;
  if (rel) {
    rmap.resize(m.n[k + 1]);
    cmap.resize(m.n[k]);
    copy_to_host(m.ctx, rmap.data(), m.interior[k + 1].p, rmap.size() * 4);
    copy_to_host(m.ctx, cmap.data(), m.interior[k].p, cmap.size() * 4);
  }
  for (int64_t s = 0; s < full.nrows; ++s) {
    const int64_t r = rel ? rmap[s] : s;
    if (r < 0) continue;
    for (int64_t p = rp[s]; p < rp[s + 1]; ++p) {
      const int64_t c = rel ? cmap[col[p]] : col[p];
      if (c < 0) continue;
      a.rows[r].push_back({(int32_t)c, (int64_t)val[p]});
    }
    std::sort(a.rows[r].begin(), a.rows[r].end());
  }
  return a;
}

// The rank of D_0 on the DOFs by union-find. Its rows are a graph's
// incidence rows (-1 at the smaller live end, +1 at the larger; one entry
// where the other end is a removed boundary vertex, none where both are), so
// its rank over every field is the vertex count minus the components that do
// not reach a removed vertex (one extra "ground" vertex stands for those).
// -1 when a row is not of that form.
int64_t graph_rank(const IntMat& d0) {
  const int64_t n = d0.ncols;
  std::vector<int64_t> parent(n + 1);
  for (int64_t v = 0; v <= n; ++v) parent[v] = v;
  auto find = [&](int64_t v) {
    while (parent[v] != v) {
      parent[v] = parent[parent[v]];
      v = parent[v];
    }
    return v;
  };
  int64_t merges = 0;
  auto unite = [&](int64_t a, int64_t b) {
    a = find(a);
    b = find(b);
    if (a != b) {
      parent[a] = b;
      ++merges;
    }
  };
  for (const auto& r : d0.rows) {
    if (r.size() == 2) {
      if (r[0].second != -1 || r[1].second != 1) return -1;
      unite(r[0].first, r[1].first);
    } else if (r.size() == 1) {
      if (r[0].second != -1 && r[0].second != 1) return -1;
      unite(r[0].first, n);
    } else if (!r.empty()) {
      return -1;
    }
  }
  return merges;  // (n + 1) - components of the graph with its ground vertex
}

}  // namespace

IntMat int_transpose(const IntMat& a) {
  IntMat t;
  t.nrows = a.ncols;
  t.ncols = a.nrows;
  t.rows.assign(t.nrows, {});
  for (int64_t r = 0; r < a.nrows; ++r)
    for (const auto& e : a.rows[r]) t.rows[e.first].push_back({(int32_t)r, e.second});
  return t;  // rows come out sorted: r runs in order
}

int64_t exact_rank(const IntMat& a) {
  // the larger of the ranks over two primes near 2^30 (exact.rs:106-115), the
  // two eliminations on two host threads
  auto second = std::async(std::launch::async, [&a] {
    Work w;
    return rank_mod(a, 1000000009, w);
  });
  Work w;
  const int64_t r1 = rank_mod(a, 1000000007, w);
  return std::max(r1, second.get());
}

namespace {

// A's columns as graph edges (tail, head), when every column is an incidence
// column: -1 at its tail and +1 at its head, or one of them alone (the other
// end is the "ground" vertex A.nrows), or empty (a loop at ground). False
// otherwise. D_0^T (the boundary d_1) has this form, absolute and relative.
bool graph_columns(const IntMat& A, std::vector<int64_t>& tail, std::vector<int64_t>& head) {
  const int64_t g = A.nrows;
  tail.assign(A.ncols, g);
  head.assign(A.ncols, g);
  std::vector<int8_t> seen(A.ncols, 0);  // bit 0: a -1 seen, bit 1: a +1 seen
  for (int64_t r = 0; r < A.nrows; ++r)
    for (const auto& e : A.rows[r]) {
      const int bit = e.second == -1 ? 1 : e.second == 1 ? 2 : 0;
      if (!bit || (seen[e.first] & bit)) return false;
      seen[e.first] |= bit;
      (bit == 1 ? tail : head)[e.first] = r;
    }
  return true;
}

// The RREF's free columns and the kernel vectors v_f of the kept ones, for a
// graph-shaped A. The pivot columns are the lowest-index column basis, i.e.
// the edges Kruskal keeps in index order (union-find); v_f is the one vector
// of ker A with a 1 at f on f and pivots before it, i.e. f's fundamental cycle
// in that forest (the forest's path between f's ends is the same in the final
// forest), +1 on the edges it traverses forward.
struct GraphKernel {
  std::vector<int64_t> tail, head, parent_edge, up, depth;
  std::vector<char> pivot;

  void build(int64_t nv) {  // nv vertices including ground
    const int64_t ne = (int64_t)tail.size();
    std::vector<int64_t> uf(nv);
    for (int64_t v = 0; v < nv; ++v) uf[v] = v;
    auto find = [&](int64_t v) {
      while (uf[v] != v) {
        uf[v] = uf[uf[v]];
        v = uf[v];
      }
      return v;
    };
    pivot.assign(ne, 0);
    std::vector<std::vector<std::pair<int64_t, int64_t>>> adj(nv);  // (neighbour, edge)
    for (int64_t e = 0; e < ne; ++e) {
      const int64_t a = find(tail[e]), b = find(head[e]);
      if (a == b) continue;
      uf[a] = b;
      pivot[e] = 1;
      adj[tail[e]].push_back({head[e], e});
      adj[head[e]].push_back({tail[e], e});
    }
    // root every tree; parent edge and depth per vertex
    parent_edge.assign(nv, -1);
    up.assign(nv, -1);
    depth.assign(nv, -1);
    std::vector<int64_t> stack;
    for (int64_t r = 0; r < nv; ++r) {
      if (depth[r] >= 0) continue;
      depth[r] = 0;
      stack.push_back(r);
      while (!stack.empty()) {
        const int64_t v = stack.back();
        stack.pop_back();
        for (const auto& [u, e] : adj[v])
          if (depth[u] < 0) {
            depth[u] = depth[v] + 1;
            up[u] = v;
            parent_edge[u] = e;
            stack.push_back(u);
          }
      }
    }
  }

  // v_f: +1 on f, then the forest path from head(f) back to tail(f)
  std::vector<Rat> cycle(int64_t f) const {
    std::vector<Rat> x(tail.size());
    x[f] = Rat{1, 1};
    if (tail[f] == head[f]) return x;  // a loop: the column is zero
    // walk x from head(f) and y from tail(f) up to their common ancestor; the
    // edge above a vertex v on the head side is traversed v -> up(v), on the
    // tail side up(v) -> v
    int64_t a = head[f], b = tail[f];
    auto mark = [&](int64_t v, bool from_v) {
      const int64_t e = parent_edge[v];
      const int64_t from = from_v ? v : up[v], to = from_v ? up[v] : v;
      x[e] = Rat{(tail[e] == from && head[e] == to) ? 1 : -1, 1};
    };
    while (depth[a] > depth[b]) {
      mark(a, true);
      a = up[a];
    }
    while (depth[b] > depth[a]) {
      mark(b, false);
      b = up[b];
    }
    while (a != b) {
      FEEC_REQUIRE(up[a] >= 0 && up[b] >= 0, FEEC_E_DEVICE, "a free edge's ends lie in different trees");
      mark(a, true);
      a = up[a];
      mark(b, false);
      b = up[b];
    }
    return x;
  }
};

}  // namespace

// Representatives of ker A / im B over Q, chosen as exact.rs:178-264 chooses
// them: the RREF kernel vector v_f of every free column f of A, in column
// order, kept when it enlarges span(im B, the v before it), each returned
// primitive. A's RREF is never formed: a vector x of ker A is fixed by its
// values on the free columns, and v_f is the one with a 1 at f and 0 at the
// other free columns, so v_f lies in that span exactly when some column of B,
// read on the free columns, can be combined to end (in free-column order) at
// f. Those ends are the pivots of B's free-row columns eliminated by their
// last entry; the kept v_f are the other free columns, and each is formed by
// one back-substitution through A's echelon rows (or, when A is a graph's
// incidence, as a fundamental cycle).
std::vector<std::vector<int64_t>> quotient_generators(const IntMat& A, const IntMat& B) {
  FEEC_REQUIRE(A.ncols == B.nrows, FEEC_E_DEVICE, "quotient_generators: the maps do not compose");
  const int64_t n = A.ncols;
  Work w;
  GraphKernel gk;
  const bool graph = graph_columns(A, gk.tail, gk.head);
  // A's echelon rows over Q, leading entry 1, keyed by their leading column
  // (general A), or the Kruskal forest (graph A)
  std::vector<int32_t> pivot_of(std::max<int64_t>(n, 1), -1);
  std::vector<Row<Rat>> piv;
  if (graph) {
    gk.build(A.nrows + 1);
    for (int64_t c = 0; c < n; ++c)
      if (gk.pivot[c]) pivot_of[c] = 0;
  } else {
    Row<Rat> row, tmp;
    for (const auto& r0 : A.rows) {
      row.clear();
      for (const auto& e : r0)
        if (e.second) row.push_back({e.first, Rat{e.second, 1}});
      while (!row.empty()) {
        const int32_t lead = row[0].first;
        const Rat c = row[0].second;
        if (pivot_of[lead] < 0) {
          for (auto& e : row) e.second = e.second / c;
          pivot_of[lead] = (int32_t)piv.size();
          piv.push_back(row);
          break;
        }
        const Row<Rat>& pv = piv[pivot_of[lead]];
        w.add((int64_t)(row.size() + pv.size()));
        row_axpy<Rat>(row, pv, tmp, [&](Rat x, Rat y) { return x - c * y; },
                      [&](Rat y) { return Rat{0, 1} - c * y; });
        row.swap(tmp);
      }
    }
  }
  // the free columns and their positions in free order
  std::vector<int32_t> free_pos(std::max<int64_t>(n, 1), -1), free_cols;
  for (int64_t c = 0; c < n; ++c)
    if (pivot_of[c] < 0) {
      free_pos[c] = (int32_t)free_cols.size();
      free_cols.push_back((int32_t)c);
    }
  const int64_t nf = (int64_t)free_cols.size();
  // B's columns on the free rows, eliminated by their last entry
  std::vector<Row<Rat>> bcols(B.ncols);
  for (int64_t r = 0; r < B.nrows; ++r) {
    const int32_t fp = free_pos[r];
    if (fp < 0) continue;
    for (const auto& e : B.rows[r])
      if (e.second) bcols[e.first].push_back({fp, Rat{e.second, 1}});
  }
  std::vector<char> ends(std::max<int64_t>(nf, 1), 0);
  {
    std::vector<int32_t> end_of(std::max<int64_t>(nf, 1), -1);
    std::vector<Row<Rat>> epiv;
    Row<Rat> tmp;
    for (auto& y : bcols) {
      while (!y.empty()) {
        const int32_t last = y.back().first;
        const Rat c = y.back().second;
        if (end_of[last] < 0) {
          for (auto& e : y) e.second = e.second / c;
          end_of[last] = (int32_t)epiv.size();
          ends[last] = 1;
          epiv.push_back(y);
          break;
        }
        const Row<Rat>& pv = epiv[end_of[last]];
        w.add((int64_t)(y.size() + pv.size()));
        row_axpy<Rat>(y, pv, tmp, [&](Rat x, Rat v) { return x - c * v; },
                      [&](Rat v) { return Rat{0, 1} - c * v; });
        y.swap(tmp);
      }
    }
  }
  // the kept kernel vectors
  std::vector<int32_t> order(piv.size());
  for (size_t i = 0; i < piv.size(); ++i) order[i] = (int32_t)i;
  std::sort(order.begin(), order.end(),
            [&](int32_t a, int32_t b) { return piv[a][0].first > piv[b][0].first; });
  std::vector<std::vector<int64_t>> out;
  for (int64_t fi = 0; fi < nf; ++fi) {
    if (ends[fi]) continue;
    if (graph) {
      out.push_back(primitive(gk.cycle(free_cols[fi])));
      continue;
    }
    std::vector<Rat> x(n);
    x[free_cols[fi]] = Rat{1, 1};
    for (int32_t i : order) {
      const Row<Rat>& row = piv[i];
      Rat s{0, 1};
      for (size_t q = 1; q < row.size(); ++q)
        if (!x[row[q].first].zero()) s = s + row[q].second * x[row[q].first];
      x[row[0].first] = Rat{0, 1} - s;
      w.add((int64_t)row.size());
    }
    out.push_back(primitive(x));
  }
  return out;
}

// b_k of the complex (FEEC_FULL) or relative to its boundary (FEEC_RELATIVE):
// n_k - rank d_k - rank d_{k+1}, with d_k = D_{k-1}^T (homology.rs:47-52, 81-98).
// A Kuhn box is a ball: b = (1, 0, ..., 0), relative (0, ..., 0, 1).
int64_t betti_number(Mesh& m, int bc, int k) {
  FEEC_REQUIRE(k >= 0 && k <= m.d, FEEC_E_ARG, "grade out of range");
  if (kuhn_grid_n(m) > 0) return (bc == FEEC_RELATIVE ? k == m.d : k == 0) ? 1 : 0;
  {
    std::lock_guard<std::mutex> lock(m.topo_mu);
    auto it = m.betti.find({bc, k});
    if (it != m.betti.end()) return it->second;
  }
  // rank d_j = rank D_{j-1}; D_0's by union-find
  auto rank_of = [&](int j, const IntMat& d) -> int64_t {
    if (j == 0) {
      const int64_t g = graph_rank(d);
      if (g >= 0) return g;
    }
    return exact_rank(int_transpose(d));
  };
  const IntMat lo = coboundary_int(m, bc, k - 1), hi = coboundary_int(m, bc, k);
  const int64_t nk = lo.nrows;
  auto r_hi = std::async(std::launch::async, [&] { return rank_of(k, hi); });
  const int64_t r_lo = rank_of(k - 1, lo);
  const int64_t b = nk - r_lo - r_hi.get();
  std::lock_guard<std::mutex> lock(m.topo_mu);
  m.betti[{bc, k}] = b;
  return b;
}

// Integral cocycles of grade k on the complex's DOFs (cohomology.rs:55-90):
// ker D_k / im D_{k-1}.
namespace {
// generators of one kind (0 cocycles, 1 cycles), computed once per mesh
std::vector<std::vector<int64_t>> cached_generators(Mesh& m, int bc, int k, int kind) {
  {
    std::lock_guard<std::mutex> lock(m.topo_mu);
    auto it = m.generators.find({bc, k, kind});
    if (it != m.generators.end()) return it->second;
  }
  std::vector<std::vector<int64_t>> g =
      kind == 0 ? quotient_generators(coboundary_int(m, bc, k), coboundary_int(m, bc, k - 1))
                : quotient_generators(int_transpose(coboundary_int(m, bc, k - 1)),
                                      int_transpose(coboundary_int(m, bc, k)));
  std::lock_guard<std::mutex> lock(m.topo_mu);
  m.generators[{bc, k, kind}] = g;
  return g;
}

// Both kinds at once: the two coboundaries are copied once, and the two
// eliminations (host-only) run on two threads.
void cache_both_generators(Mesh& m, int bc, int k) {
  {
    std::lock_guard<std::mutex> lock(m.topo_mu);
    if (m.generators.count({bc, k, 0}) && m.generators.count({bc, k, 1})) return;
  }
  const IntMat lo = coboundary_int(m, bc, k - 1), hi = coboundary_int(m, bc, k);
  auto cyc = std::async(std::launch::async,
                        [&] { return quotient_generators(int_transpose(lo), int_transpose(hi)); });
  std::vector<std::vector<int64_t>> co = quotient_generators(hi, lo);
  std::vector<std::vector<int64_t>> ho = cyc.get();
  std::lock_guard<std::mutex> lock(m.topo_mu);
  m.generators[{bc, k, 0}] = std::move(co);
  m.generators[{bc, k, 1}] = std::move(ho);
}
}  // namespace

std::vector<std::vector<int64_t>> cohomology_generators(Mesh& m, int bc, int k) {
  FEEC_REQUIRE(k >= 0 && k <= m.d, FEEC_E_ARG, "grade out of range");
  return cached_generators(m, bc, k, 0);
}

// Integral cycles of grade k on the complex's DOFs (homology.rs:116-153):
// ker d_k / im d_{k+1}, d_k = D_{k-1}^T.
std::vector<std::vector<int64_t>> homology_generators(Mesh& m, int bc, int k) {
  FEEC_REQUIRE(k >= 0 && k <= m.d, FEEC_E_ARG, "grade out of range");
  return cached_generators(m, bc, k, 1);
}

// The harmonic bases from integral cohomology (harmonic.rs:118-175): each
// cocycle z projected to h = z - D p with D^T M D p = D^T M z (CG, identity
// preconditioner, rtol 1e-12; D = D_{k-1} on the DOFs), period-normalized
// against the integral cycles (integral = H (P^T)^-1, P_ij = <z_i, c_j>), and
// made M-orthonormal through the Cholesky factor of its Gram matrix
// (orthonormal = integral L^-T). Both [count][ndofs].
void harmonics_exact(Geom& g, int bc, int k, std::vector<double>& integral, std::vector<double>& orthonormal,
                     int64_t& count) {
  Mesh& m = *g.mesh;
  Ctx* ctx = m.ctx;
  cache_both_generators(m, bc, k);
  const std::vector<std::vector<int64_t>> z = cohomology_generators(m, bc, k);
  const std::vector<std::vector<int64_t>> cyc = homology_generators(m, bc, k);
  const int64_t b = (int64_t)z.size();
  FEEC_REQUIRE((int64_t)cyc.size() == b, FEEC_E_DEVICE, "cocycle and cycle counts differ");
  count = b;
  integral.clear();
  orthonormal.clear();
  if (!b) return;
  FEEC_REQUIRE(b <= kMaxHarmonics, FEEC_E_UNSUPPORTED, "more harmonic forms than the KKT border holds (8)");
  Csr M;
  assemble_form(g, bc, FEEC_MASS, k, M);
  const int64_t n = M.nrows;
  FEEC_REQUIRE((int64_t)z[0].size() == n, FEEC_E_DEVICE, "cocycle length differs from the DOF count");
  // D = D_{k-1} on the DOFs and the normal operator D^T M D
  Csr D, Dt, MD, N;
  const bool has_d = k >= 1;
  if (has_d) {
    if (bc == FEEC_RELATIVE) {
      Csr full;
      mesh_coboundary(m, k - 1, full);
      m.ensure_interior();
      csr_select(full, m.interior[k].p, m.n_interior[k], m.interior[k - 1].p, m.n_interior[k - 1], D);
    } else {
      mesh_coboundary(m, k - 1, D);
    }
    csr_transpose(D, Dt);
    csr_mul(M, D, MD);
    csr_mul(Dt, MD, N);
  }
  const int64_t np = has_d ? D.ncols : 0;
  std::vector<double> H((size_t)b * n);
  DBuf<double> dz(ctx, std::max<int64_t>(n, 1)), mz(ctx, std::max<int64_t>(n, 1)),
      rhs(ctx, std::max<int64_t>(np, 1)), p(ctx, std::max<int64_t>(np, 1)), one(ctx, std::max<int64_t>(np, 1)),
      dp(ctx, std::max<int64_t>(n, 1));
  if (np) {
    std::vector<double> ones(np, 1.0);
    copy_to_device(ctx, one.p, ones.data(), np * 8);
  }
  for (int64_t i = 0; i < b; ++i) {
    std::vector<double> zi(n);
    for (int64_t j = 0; j < n; ++j) zi[j] = (double)z[i][j];
    double* hi = &H[(size_t)i * n];
    std::copy(zi.begin(), zi.end(), hi);
    if (!np) continue;
    copy_to_device(ctx, dz.p, zi.data(), n * 8);
    spmv64(ctx, M, dz.p, mz.p);
    spmv64(ctx, Dt, mz.p, rhs.p);
    FEEC_HIP(hipMemsetAsync(p.p, 0, np * 8, ctx->stream));
    std::vector<double> hr(np);
    copy_to_host(ctx, hr.data(), rhs.p, np * 8);
    double rn = 0.0;
    for (double v : hr) rn += v * v;
    if (rn > 0.0) {
      const KrylovStats ks = cg_solve_device(ctx, N, one.p, rhs.p, p.p, 1e-12, std::max<int64_t>(10 * np, 1000));
      FEEC_REQUIRE(ks.converged, FEEC_E_NOT_CONVERGED, "the harmonic projection's CG did not converge");
    }
    spmv64(ctx, D, p.p, dp.p);
    std::vector<double> hd(n);
    copy_to_host(ctx, hd.data(), dp.p, n * 8);
    for (int64_t j = 0; j < n; ++j) hi[j] -= hd[j];
  }
  // period normalization: integral = H (P^T)^-1, P_ij = <z_i, c_j>
  std::vector<double> Pt((size_t)b * b), inv((size_t)b * b, 0.0);
  for (int64_t i = 0; i < b; ++i)
    for (int64_t j = 0; j < b; ++j) {
      int64_t s = 0;
      for (int64_t q = 0; q < n; ++q) s += z[i][q] * cyc[j][q];
      Pt[j * b + i] = (double)s;  // (P^T)_{ji}
    }
  for (int64_t i = 0; i < b; ++i) inv[i * b + i] = 1.0;
  for (int64_t c = 0; c < b; ++c) {  // Gauss-Jordan with partial pivoting
    int64_t piv = c;
    for (int64_t r = c + 1; r < b; ++r)
      if (std::abs(Pt[r * b + c]) > std::abs(Pt[piv * b + c])) piv = r;
    FEEC_REQUIRE(Pt[piv * b + c] != 0.0, FEEC_E_DEVICE, "the period matrix is singular");
    for (int64_t q = 0; q < b; ++q) {
      std::swap(Pt[c * b + q], Pt[piv * b + q]);
      std::swap(inv[c * b + q], inv[piv * b + q]);
    }
    const double d = Pt[c * b + c];
    for (int64_t q = 0; q < b; ++q) {
      Pt[c * b + q] /= d;
      inv[c * b + q] /= d;
    }
    for (int64_t r = 0; r < b; ++r) {
      if (r == c || Pt[r * b + c] == 0.0) continue;
      const double f = Pt[r * b + c];
      for (int64_t q = 0; q < b; ++q) {
        Pt[r * b + q] -= f * Pt[c * b + q];
        inv[r * b + q] -= f * inv[c * b + q];
      }
    }
  }
  // integral column j = sum_i H column i * inv(i, j)
  integral.assign((size_t)b * n, 0.0);
  for (int64_t j = 0; j < b; ++j)
    for (int64_t i = 0; i < b; ++i) {
      const double f = inv[i * b + j];
      if (f == 0.0) continue;
      for (int64_t q = 0; q < n; ++q) integral[(size_t)j * n + q] += H[(size_t)i * n + q] * f;
    }
  // Gram matrix in M and its Cholesky factor
  std::vector<double> G((size_t)b * b), Mi((size_t)b * n);
  for (int64_t j = 0; j < b; ++j) {
    copy_to_device(ctx, dz.p, &integral[(size_t)j * n], n * 8);
    spmv64(ctx, M, dz.p, mz.p);
    copy_to_host(ctx, &Mi[(size_t)j * n], mz.p, n * 8);
  }
  for (int64_t i = 0; i < b; ++i)
    for (int64_t j = 0; j < b; ++j) {
      double s = 0.0;
      for (int64_t q = 0; q < n; ++q) s += integral[(size_t)i * n + q] * Mi[(size_t)j * n + q];
      G[i * b + j] = s;
    }
  std::vector<double> L((size_t)b * b, 0.0);
  for (int64_t j = 0; j < b; ++j) {
    double s = G[j * b + j];
    for (int64_t q = 0; q < j; ++q) s -= L[j * b + q] * L[j * b + q];
    FEEC_REQUIRE(s > 0.0, FEEC_E_UNSUPPORTED,
                 "the L2 pairing is not definite on the harmonic space (the reference falls back to an "
                 "eigensolve there, which is not built)");
    L[j * b + j] = std::sqrt(s);
    for (int64_t i = j + 1; i < b; ++i) {
      double t = G[i * b + j];
      for (int64_t q = 0; q < j; ++q) t -= L[i * b + q] * L[j * b + q];
      L[i * b + j] = t / L[j * b + j];
    }
  }
  // orthonormal = integral L^-T: row q of the result solves L y = integral row q
  orthonormal.assign((size_t)b * n, 0.0);
  for (int64_t q = 0; q < n; ++q)
    for (int64_t i = 0; i < b; ++i) {
      double t = integral[(size_t)i * n + q];
      for (int64_t j = 0; j < i; ++j) t -= L[i * b + j] * orthonormal[(size_t)j * n + q];
      orthonormal[(size_t)i * n + q] = t / L[i * b + i];
    }
}

}  // namespace feec
