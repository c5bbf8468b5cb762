// Exercises the host-only code of libfeec (the meshers, option defaults,
// status strings, element shapes, argument checks, the multigrid set-up's
// dense Cholesky inverse and tridiagonal eigenvalue bisection, and the exact
// (co)homology eliminations) in a build
// instrumented by AddressSanitizer and UndefinedBehaviorSanitizer. No GPU is
// touched. Exit status 0 and the line "host_check ok" mean every check held.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "feec.h"
#include "../../formoniq_amd/csrc/feec_internal.hpp"

static int failures = 0;
#define EXPECT(c)                                                     \
  do {                                                                \
    if (!(c)) {                                                       \
      std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c); \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

static void meshers() {
  const int sizes[5] = {0, 7, 5, 3, 2};
  for (int d = 1; d <= 4; ++d) {
    int64_t nc = 0, nv = 0;
    EXPECT(feec_kuhn_counts(d, sizes[d], &nc, &nv) == FEEC_OK);
    std::vector<int64_t> cells((size_t)nc * (d + 1));
    std::vector<double> coords((size_t)nv * d);
    EXPECT(feec_kuhn_grid(d, sizes[d], cells.data(), coords.data()) == FEEC_OK);
    bool in_range = true;
    for (int64_t v : cells) in_range &= v >= 0 && v < nv;
    EXPECT(in_range);
    // each cell's vertices step along distinct axes: a Kuhn simplex
    EXPECT(feec_jitter_interior(d, sizes[d], coords.data(), 0.2, 7, 0xFu) == FEEC_OK);
    for (double c : coords) EXPECT(c >= -0.25 && c <= 1.25);
    EXPECT(feec_kuhn_grid(d, sizes[d], nullptr, coords.data()) == FEEC_OK);
    EXPECT(feec_kuhn_grid(d, sizes[d], cells.data(), nullptr) == FEEC_OK);
  }
  int64_t nc = 0, nv = 0;
  EXPECT(feec_kuhn_counts(5, 2, &nc, &nv) == FEEC_E_ARG);
  EXPECT(std::strlen(feec_last_error()) > 0);
  EXPECT(feec_kuhn_counts(2, 0, &nc, &nv) == FEEC_E_ARG);
  EXPECT(feec_jitter_interior(2, 3, nullptr, 0.1, 1, 3u) == FEEC_E_ARG);
  for (int64_t n : {0, 1, 2, 7, 1000}) {
    std::vector<double> x((size_t)n + 1, 12345.0);
    EXPECT(feec_normal_vector(n, 42, n ? x.data() : nullptr) == FEEC_OK);
    for (int64_t i = 0; i < n; ++i) EXPECT(std::isfinite(x[i]));
    EXPECT(x[n] == 12345.0);  // nothing written past n
  }
  EXPECT(feec_normal_vector(-1, 0, nullptr) == FEEC_E_ARG);
}

static void options_and_strings() {
  feec_solve_opts o;
  std::memset(&o, 0xA5, sizeof(o));
  feec_solve_opts_default(&o);
  EXPECT(o.rtol > 0.0 && o.rtol < 1.0);
  EXPECT(o.precision == FEEC_PREC_FP64_PC32);
  EXPECT(o.preconditioner == FEEC_PC_AUTO);
  // values outside the enumerations, as a C or Rust caller may pass them
  for (int s : {0, 1, 5, 10, 11, 99, -3}) {
    const char* t = feec_status_string((feec_status)s);
    EXPECT(t != nullptr && std::strlen(t) > 0);
  }
  EXPECT(std::strcmp(feec_status_string((feec_status)99), "unknown") == 0);
  for (int d = 0; d <= 5; ++d)
    for (int kind = FEEC_MASS; kind <= FEEC_DIF_BOTH; ++kind)
      for (int k = -1; k <= d + 1; ++k) {
        int r = -1, c = -1;
        EXPECT(feec_element_shape(d, (feec_form_kind)kind, k, &r, &c) == FEEC_OK);
        EXPECT(r >= 0 && c >= 0);
      }
  int r, c;
  EXPECT(feec_element_shape(3, (feec_form_kind)99, 1, &r, &c) == FEEC_E_ARG);
  EXPECT(feec_element_shape(3, (feec_form_kind)-1, 1, &r, &c) == FEEC_E_ARG);
  EXPECT(feec_element_shape(3, FEEC_MASS, 1, nullptr, &c) == FEEC_E_ARG);
  EXPECT(feec_assemble(nullptr, (feec_bc)7, FEEC_MASS, 1, nullptr) == FEEC_E_ARG);
}

static void dense_inverse() {
  // SPD: A = L L^T + n I from a deterministic L; check A * inv(A) = I
  for (int n : {1, 2, 5, 17, 64}) {
    std::vector<double> a((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double s = i == j ? n : 0.0;
        for (int k = 0; k <= std::min(i, j); ++k) s += std::sin(1.0 + i + 3.0 * k) * std::sin(1.0 + j + 3.0 * k);
        a[(size_t)i * n + j] = s;
      }
    std::vector<double> inv = a;
    EXPECT(feec::spd_inverse(n, inv));
    double err = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += a[(size_t)i * n + k] * inv[(size_t)k * n + j];
        err = std::max(err, std::fabs(s - (i == j ? 1.0 : 0.0)));
      }
    EXPECT(err < 1e-10);
  }
  std::vector<double> indef = {1.0, 2.0, 2.0, 1.0};  // eigenvalues 3, -1
  EXPECT(!feec::spd_inverse(2, indef));
  std::vector<double> empty;
  EXPECT(feec::spd_inverse(0, empty));
}

static void tridiagonal() {
  // 1D Laplacian tridiag(-1, 2, -1) of size n: largest eigenvalue 2 - 2 cos(n pi / (n+1))
  for (int n : {1, 2, 3, 6, 16}) {
    std::vector<double> dg(n, 2.0), off(n > 0 ? n - 1 : 0, -1.0);
    const double want = 2.0 - 2.0 * std::cos(n * M_PI / (n + 1));
    EXPECT(std::fabs(feec::tridiagonal_lmax(dg, off) - want) < 1e-12 * want);
  }
  // a diagonal matrix with a zero off-diagonal: the largest diagonal entry
  std::vector<double> dg = {0.5, 3.0, 1.0}, off = {0.0, 0.0};
  EXPECT(std::fabs(feec::tridiagonal_lmax(dg, off) - 3.0) < 1e-12);
}

static feec::IntMat dense_int(int64_t r, int64_t c, const std::vector<int64_t>& v) {
  feec::IntMat a;
  a.nrows = r;
  a.ncols = c;
  a.rows.assign(r, {});
  for (int64_t i = 0; i < r; ++i)
    for (int64_t j = 0; j < c; ++j)
      if (v[i * c + j]) a.rows[i].push_back({(int32_t)j, v[i * c + j]});
  return a;
}

// The exact (co)homology routines on incidence matrices built here: the
// triangle's boundary (a circle), a long path (a tree) and a long cycle.
static void exact_topology() {
  using feec::IntMat;
  const IntMat d0 = dense_int(3, 3, {-1, 1, 0, -1, 0, 1, 0, -1, 1});  // edges 01, 02, 12
  const IntMat no_rows = dense_int(0, 3, {}), no_cols = dense_int(3, 0, {});
  EXPECT(feec::exact_rank(d0) == 2);
  EXPECT(feec::exact_rank(no_rows) == 0);
  const auto h0 = feec::quotient_generators(d0, no_cols);  // H^0: the constant
  EXPECT(h0.size() == 1 && h0[0] == std::vector<int64_t>({1, 1, 1}));
  const auto h1 = feec::quotient_generators(no_rows, d0);  // H^1: the first edge's cochain
  EXPECT(h1.size() == 1 && h1[0] == std::vector<int64_t>({1, 0, 0}));
  const auto c1 = feec::quotient_generators(feec::int_transpose(d0), no_cols);  // H_1: the loop
  EXPECT(c1.size() == 1 && c1[0] == std::vector<int64_t>({1, -1, 1}));
  const auto c0 = feec::quotient_generators(no_rows, feec::int_transpose(d0));  // H_0: a vertex
  EXPECT(c0.size() == 1 && c0[0] == std::vector<int64_t>({1, 0, 0}));
  for (bool closed : {false, true}) {
    const int64_t n = 3000, ne = closed ? n : n - 1;
    IntMat d;  // edge i joins vertices i and i + 1 (mod n)
    d.nrows = ne;
    d.ncols = n;
    d.rows.assign(ne, {});
    for (int64_t e = 0; e < ne; ++e) {
      const int32_t a = (int32_t)e, b = (int32_t)((e + 1) % n);
      if (a < b)
        d.rows[e] = {{a, -1}, {b, 1}};
      else
        d.rows[e] = {{b, 1}, {a, -1}};
    }
    EXPECT(feec::exact_rank(d) == n - 1);
    const auto z0 = feec::quotient_generators(d, dense_int(n, 0, {}));
    EXPECT(z0.size() == 1 && z0[0] == std::vector<int64_t>(n, 1));
    const auto cyc = feec::quotient_generators(feec::int_transpose(d), dense_int(ne, 0, {}));
    EXPECT(cyc.size() == (closed ? 1u : 0u));
    if (closed && cyc.size() == 1) {
      bool unit = true;
      for (int64_t v : cyc[0]) unit = unit && (v == 1 || v == -1);
      EXPECT(unit && cyc[0][0] == 1);
    }
  }
}

// The ingestion readers on valid files, on every truncation of them and on
// byte-flipped copies: each call returns a status, never reads out of bounds.
static void ingestion() {
  const char* square =
      "$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1 4 1 4\n2 1 0 4\n1\n2\n3\n4\n"
      "0 0 0\n1 0 0\n0 1 0\n1 1 0\n$EndNodes\n$Elements\n1 2 1 2\n2 1 2 2\n1 1 2 3\n"
      "2 2 4 3\n$EndElements\n";
  const int64_t n = (int64_t)std::strlen(square);
  int dim = -1, cd = -1;
  int64_t nv = 0, nc = 0;
  EXPECT(feec_gmsh_read(square, n, &dim, &cd, &nv, &nc, nullptr, nullptr) == FEEC_OK);
  EXPECT(dim == 2 && cd == 2 && nv == 4 && nc == 2);
  std::vector<int64_t> cells(nc * (dim + 1));
  std::vector<double> coords(nv * cd);
  EXPECT(feec_gmsh_read(square, n, &dim, &cd, &nv, &nc, cells.data(), coords.data()) == FEEC_OK);
  EXPECT(cells[3] == 1 && cells[4] == 3 && cells[5] == 2);
  for (int64_t cut = 0; cut < n; ++cut) {
    std::vector<char> part(square, square + cut);  // exactly cut bytes, no terminator
    feec_gmsh_read(part.data(), cut, &dim, &cd, &nv, &nc, nullptr, nullptr);
  }
  uint64_t state = 12345;
  auto next = [&] {
    state = state * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t)(state >> 33);
  };
  for (int trial = 0; trial < 2000; ++trial) {
    std::vector<char> m(square, square + n);
    m[next() % n] = (char)(next() & 0xff);
    if (feec_gmsh_read(m.data(), n, &dim, &cd, &nv, &nc, nullptr, nullptr) == FEEC_OK) {
      std::vector<int64_t> c2(nc * (dim + 1));
      std::vector<double> x2(nv * cd);
      EXPECT(feec_gmsh_read(m.data(), n, &dim, &cd, &nv, &nc, c2.data(), x2.data()) == FEEC_OK);
    }
  }
  // CBOR: a skeleton and a length vector written, then read whole, cut and flipped
  const int64_t kc[6] = {0, 1, 2, 1, 2, 300};
  int64_t len = 0;
  EXPECT(feec_cbor_write_cells(2, 2, kc, nullptr, &len) == FEEC_OK);
  std::vector<uint8_t> cb(len);
  EXPECT(feec_cbor_write_cells(2, 2, kc, cb.data(), &len) == FEEC_OK);
  const double lv[5] = {1.0, 0.1, 65504.0, 1e300, -0.0};
  int64_t llen = 0;
  EXPECT(feec_cbor_write_lengths(5, lv, nullptr, &llen) == FEEC_OK);
  std::vector<uint8_t> lb(llen);
  EXPECT(feec_cbor_write_lengths(5, lv, lb.data(), &llen) == FEEC_OK);
  int64_t ne = 0;
  std::vector<double> back(5);
  EXPECT(feec_cbor_read_lengths(lb.data(), llen, &ne, back.data()) == FEEC_OK && ne == 5);
  for (int i = 0; i < 5; ++i) EXPECT(back[i] == lv[i]);
  for (const std::vector<uint8_t>* buf : {&cb, &lb}) {
    for (size_t cut = 0; cut < buf->size(); ++cut) {
      std::vector<uint8_t> part(buf->begin(), buf->begin() + cut);
      feec_cbor_read_cells(part.data(), (int64_t)cut, &dim, &nc, nullptr);
      feec_cbor_read_lengths(part.data(), (int64_t)cut, &ne, nullptr);
    }
    for (int trial = 0; trial < 2000; ++trial) {
      std::vector<uint8_t> m(*buf);
      m[next() % m.size()] = (uint8_t)next();
      if (feec_cbor_read_cells(m.data(), (int64_t)m.size(), &dim, &nc, nullptr) == FEEC_OK) {
        std::vector<int64_t> c2(nc * (dim + 1));
        feec_cbor_read_cells(m.data(), (int64_t)m.size(), &dim, &nc, c2.data());
      }
      if (feec_cbor_read_lengths(m.data(), (int64_t)m.size(), &ne, nullptr) == FEEC_OK) {
        std::vector<double> v2(ne);
        feec_cbor_read_lengths(m.data(), (int64_t)m.size(), &ne, v2.data());
      }
    }
  }
}

int main() {
  exact_topology();
  ingestion();
  meshers();
  options_and_strings();
  dense_inverse();
  tridiagonal();
  if (failures) {
    std::fprintf(stderr, "host_check: %d failures\n", failures);
    return 1;
  }
  std::printf("host_check ok\n");
  return 0;
}
