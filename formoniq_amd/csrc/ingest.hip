// Request ingestion: meshes and geometries as the reference stores and reads
// them. Host code only.
//
// Reference behaviour (luiswirth/formoniq, crates/):
//   gmsh2coord_cells_ordered           regge/src/io/gmsh.rs:84-175 (mshio 0.4.2)
//   VertexRelabelling::of_used         simplicial/src/topology/relabel.rs:28-55
//   save_cbor / load_cbor              simplicial/src/io/cbor.rs:14-22 (ciborium 0.2.2)
//   Complex::{save, load}              simplicial/src/topology/complex.rs:91-99
//   Skeleton serialization             simplicial/src/topology/skeleton.rs:29-47
//   MeshLengthsSq::{save, load}        regge/src/lengths/mesh.rs:244-251
//
// gmsh: the ASCII form of MSH 4.1. Nodes are indexed by tag - 1 in the order
// the file lists them (the reference's convention); the cells are the elements
// of the highest simplex dimension present (points, lines, triangles,
// tetrahedra; every other element type is skipped); the vertices the cells
// name are renumbered by rank, and their coordinates kept in that order; the
// third coordinate is dropped when it is 0 on every node.
//
// CBOR (RFC 8949). A Complex is stored as its cell skeleton: an array of
// simplices, each a map {"vertices": [v0, v1, ...]} (serde's struct encoding
// through ciborium). A MeshLengthsSq is a map {"vector": V} with V nalgebra's
// DVector encoding, its VecStorage {"data": [...], "nrows": n, "ncols": null}.
// The reader accepts every integer and float width and definite or indefinite
// lengths, and also takes V as a plain array of numbers. The writer emits the
// encodings above with each float in the shortest of binary16/32/64 that holds
// it exactly, as ciborium does.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "feec_internal.hpp"

namespace feec {
namespace {

// ---------------------------------------------------------------------------
// gmsh MSH 4.1 (ASCII)
// ---------------------------------------------------------------------------
struct Lines {
  const char* p;
  const char* end;
  // next whitespace-separated token; false at the end of the input
  bool token(std::string& t) {
    while (p < end && (unsigned char)*p <= ' ') ++p;
    if (p >= end) return false;
    const char* b = p;
    while (p < end && (unsigned char)*p > ' ') ++p;
    t.assign(b, p);
    return true;
  }
  std::string need() {
    std::string t;
    FEEC_REQUIRE(token(t), FEEC_E_ARG, "malformed .msh: the file ends inside a section");
    return t;
  }
  int64_t integer() {
    const std::string t = need();
    char* e = nullptr;
    const long long v = std::strtoll(t.c_str(), &e, 10);
    FEEC_REQUIRE(e && *e == 0, FEEC_E_ARG, "malformed .msh: expected an integer, got '" + t + "'");
    return v;
  }
  double real() {
    const std::string t = need();
    char* e = nullptr;
    const double v = std::strtod(t.c_str(), &e);
    FEEC_REQUIRE(e && *e == 0, FEEC_E_ARG, "malformed .msh: expected a number, got '" + t + "'");
    return v;
  }
  // skip to the token `end_tag` (a section's $End... line)
  void skip_to(const std::string& end_tag) {
    std::string t;
    while (token(t))
      if (t == end_tag) return;
    fail(FEEC_E_ARG, "malformed .msh: no " + end_tag);
  }
};

struct Gmsh {
  int dim = 0, coord_dim = 3;
  std::vector<int64_t> words;  // [ncells][dim+1], file order, renumbered
  std::vector<double> coords;  // [nverts][coord_dim]
  int64_t ncells = 0, nverts = 0;
};

int simplex_dim_of(int64_t element_type) {
  switch (element_type) {
    case 15: return 0;  // 1-node point
    case 1: return 1;   // 2-node line
    case 2: return 2;   // 3-node triangle
    case 4: return 3;   // 4-node tetrahedron
  }
  return -1;
}

// Node count of a gmsh element type, for skipping the types that are not
// simplices of the lowest order (0 for a type this reader does not know).
int element_nodes(int64_t t) {
  static const int n[] = {0, 2, 3, 4, 4, 8, 6, 5, 3, 6, 9, 10, 27, 18, 14, 1, 8, 20, 15, 13};
  return (t >= 0 && t < (int64_t)(sizeof(n) / sizeof(n[0]))) ? n[t] : 0;
}

Gmsh read_gmsh(const char* bytes, int64_t len) {
  FEEC_REQUIRE(bytes || len == 0, FEEC_E_ARG, "null bytes");
  Lines in{bytes, bytes + len};
  std::string t;
  bool have_format = false, have_nodes = false, have_elements = false;
  std::vector<double> xyz;                 // nodes in file order
  std::vector<std::vector<int64_t>> by_dim[4];
  while (in.token(t)) {
    if (t == "$MeshFormat") {
      const std::string version = in.need();
      const int64_t file_type = in.integer();
      in.integer();  // data size
      FEEC_REQUIRE(version.rfind("4.1", 0) == 0, FEEC_E_UNSUPPORTED,
                   "only MSH 4.1 is read (the file says " + version + ")");
      FEEC_REQUIRE(file_type == 0, FEEC_E_UNSUPPORTED, "binary .msh is not read; write ASCII");
      in.skip_to("$EndMeshFormat");
      have_format = true;
    } else if (t == "$Nodes") {
      FEEC_REQUIRE(have_format, FEEC_E_ARG, "malformed .msh: $Nodes before $MeshFormat");
      const int64_t nblocks = in.integer(), nnodes = in.integer();
      in.integer();
      in.integer();
      // every node takes at least 7 bytes of text, so a count beyond len / 7 is malformed
      FEEC_REQUIRE(nblocks >= 0 && nnodes >= 0 && nnodes <= len / 7, FEEC_E_ARG,
                   "malformed .msh: node counts");
      xyz.reserve(3 * nnodes);
      int64_t seen = 0;
      for (int64_t b = 0; b < nblocks; ++b) {
        in.integer();  // entity dim
        in.integer();  // entity tag
        const int64_t parametric = in.integer();
        const int64_t count = in.integer();
        FEEC_REQUIRE(parametric == 0, FEEC_E_UNSUPPORTED, "parametric node blocks are not read");
        FEEC_REQUIRE(count >= 0 && seen + count <= nnodes, FEEC_E_ARG,
                     "malformed .msh: node block larger than the node count");
        for (int64_t i = 0; i < count; ++i) in.integer();  // tags (tag - 1 indexes, as listed)
        for (int64_t i = 0; i < 3 * count; ++i) xyz.push_back(in.real());
        seen += count;
      }
      FEEC_REQUIRE(seen == nnodes, FEEC_E_ARG, "malformed .msh: fewer nodes than declared");
      in.skip_to("$EndNodes");
      have_nodes = true;
    } else if (t == "$Elements") {
      FEEC_REQUIRE(have_format, FEEC_E_ARG, "malformed .msh: $Elements before $MeshFormat");
      const int64_t nblocks = in.integer(), nelem = in.integer();
      in.integer();
      in.integer();
      FEEC_REQUIRE(nblocks >= 0 && nelem >= 0, FEEC_E_ARG, "malformed .msh: element counts");
      for (int64_t b = 0; b < nblocks; ++b) {
        in.integer();  // entity dim
        in.integer();  // entity tag
        const int64_t type = in.integer();
        const int64_t count = in.integer();
        FEEC_REQUIRE(count >= 0 && count <= nelem && count <= len / 4, FEEC_E_ARG,
                     "malformed .msh: element block");
        const int sd = simplex_dim_of(type);
        const int nn = sd >= 0 ? sd + 1 : element_nodes(type);
        FEEC_REQUIRE(nn > 0, FEEC_E_UNSUPPORTED,
                     "unknown gmsh element type " + std::to_string(type));
        for (int64_t e = 0; e < count; ++e) {
          in.integer();  // element tag
          std::vector<int64_t> w(nn);
          for (int i = 0; i < nn; ++i) {
            const int64_t tag = in.integer();
            FEEC_REQUIRE(tag >= 1, FEEC_E_ARG, "malformed .msh: node tags start at 1");
            w[i] = tag - 1;
          }
          if (sd >= 0) by_dim[sd].push_back(std::move(w));
        }
      }
      in.skip_to("$EndElements");
      have_elements = true;
    } else if (t.size() > 1 && t[0] == '$' && t.compare(0, 4, "$End") != 0) {
      in.skip_to("$End" + t.substr(1));  // a section this reader does not use
    } else {
      fail(FEEC_E_ARG, "malformed .msh: unexpected '" + t.substr(0, 32) + "'");
    }
  }
  FEEC_REQUIRE(have_format, FEEC_E_ARG, "malformed .msh: no $MeshFormat");
  FEEC_REQUIRE(have_nodes, FEEC_E_ARG, "no node section");
  FEEC_REQUIRE(have_elements, FEEC_E_ARG, "no element section");
  int dim = 3;
  while (dim >= 0 && by_dim[dim].empty()) --dim;
  FEEC_REQUIRE(dim >= 0, FEEC_E_ARG, "no point, line, triangle or tetrahedron elements");
  const std::vector<std::vector<int64_t>>& cells = by_dim[dim];
  const int64_t nnodes = (int64_t)xyz.size() / 3;
  std::vector<int64_t> used;
  for (const auto& w : cells)
    for (int64_t v : w) {
      FEEC_REQUIRE(v >= 0 && v < nnodes, FEEC_E_ARG, "an element names a node the file lacks");
      used.push_back(v);
    }
  std::sort(used.begin(), used.end());
  used.erase(std::unique(used.begin(), used.end()), used.end());
  Gmsh g;
  g.dim = dim;
  g.ncells = (int64_t)cells.size();
  g.nverts = (int64_t)used.size();
  g.words.reserve(cells.size() * (dim + 1));
  for (const auto& w : cells)
    for (int64_t v : w)
      g.words.push_back(std::lower_bound(used.begin(), used.end(), v) - used.begin());
  // Gmsh writes three coordinates whatever the dimension: z = 0 on every node
  // is that padding and is dropped (the reference asks only about z)
  bool flat = true;
  for (int64_t i = 0; i < nnodes; ++i) flat = flat && xyz[3 * i + 2] == 0.0;
  g.coord_dim = flat ? 2 : 3;
  g.coords.reserve(used.size() * g.coord_dim);
  for (int64_t v : used)
    for (int a = 0; a < g.coord_dim; ++a) g.coords.push_back(xyz[3 * v + a]);
  return g;
}

// ---------------------------------------------------------------------------
// CBOR (RFC 8949): a small reader of the data model and a writer
// ---------------------------------------------------------------------------
struct Cbor {
  enum Kind { UINT, NINT, BYTES, TEXT, ARRAY, MAP, FLOAT, SIMPLE } kind = SIMPLE;
  uint64_t u = 0;  // UINT / NINT (-1 - u) / SIMPLE value
  double f = 0.0;
  std::string s;                // TEXT / BYTES
  std::vector<Cbor> items;      // ARRAY; MAP as key, value, key, value, ...
};

double half_to_double(uint16_t h) {
  const int e = (h >> 10) & 0x1f, m = h & 0x3ff;
  double v;
  if (e == 0)
    v = std::ldexp((double)m, -24);
  else if (e == 31)
    v = m ? NAN : INFINITY;
  else
    v = std::ldexp((double)(m + 1024), e - 25);
  return (h & 0x8000) ? -v : v;
}

struct CborReader {
  const uint8_t* p;
  const uint8_t* end;
  int depth = 0;
  uint8_t byte() {
    FEEC_REQUIRE(p < end, FEEC_E_ARG, "malformed CBOR: truncated");
    return *p++;
  }
  uint64_t be(int n) {
    uint64_t v = 0;
    for (int i = 0; i < n; ++i) v = (v << 8) | byte();
    return v;
  }
  // argument of an initial byte; indefinite = true for additional info 31
  uint64_t arg(uint8_t ai, bool& indefinite) {
    indefinite = false;
    if (ai < 24) return ai;
    if (ai == 24) return be(1);
    if (ai == 25) return be(2);
    if (ai == 26) return be(4);
    if (ai == 27) return be(8);
    FEEC_REQUIRE(ai == 31, FEEC_E_ARG, "malformed CBOR: reserved additional information");
    indefinite = true;
    return 0;
  }
  bool at_break() {
    FEEC_REQUIRE(p < end, FEEC_E_ARG, "malformed CBOR: truncated");
    if (*p == 0xff) {
      ++p;
      return true;
    }
    return false;
  }
  Cbor item() {
    FEEC_REQUIRE(++depth <= 64, FEEC_E_ARG, "malformed CBOR: nested too deeply");
    Cbor c;
    const uint8_t ib = byte();
    const int major = ib >> 5;
    const uint8_t ai = ib & 0x1f;
    bool indef = false;
    switch (major) {
      case 0:
      case 1:
        c.kind = major == 0 ? Cbor::UINT : Cbor::NINT;
        c.u = arg(ai, indef);
        FEEC_REQUIRE(!indef, FEEC_E_ARG, "malformed CBOR: indefinite integer");
        break;
      case 2:
      case 3: {
        c.kind = major == 2 ? Cbor::BYTES : Cbor::TEXT;
        const uint64_t n = arg(ai, indef);
        if (indef) {
          while (!at_break()) {
            Cbor chunk = item();
            FEEC_REQUIRE(chunk.kind == c.kind, FEEC_E_ARG, "malformed CBOR: string chunk type");
            c.s += chunk.s;
          }
        } else {
          FEEC_REQUIRE(n <= (uint64_t)(end - p), FEEC_E_ARG, "malformed CBOR: truncated string");
          c.s.assign((const char*)p, (size_t)n);
          p += n;
        }
        break;
      }
      case 4:
      case 5: {
        c.kind = major == 4 ? Cbor::ARRAY : Cbor::MAP;
        const uint64_t n = arg(ai, indef);
        if (indef) {
          while (!at_break()) {
            c.items.push_back(item());
            if (major == 5) c.items.push_back(item());
          }
        } else {
          // every item takes at least one byte: a count beyond the input is malformed
          FEEC_REQUIRE(n <= (uint64_t)(end - p), FEEC_E_ARG, "malformed CBOR: truncated container");
          c.items.reserve((size_t)n * (major == 5 ? 2 : 1));
          for (uint64_t i = 0; i < n; ++i) {
            c.items.push_back(item());
            if (major == 5) c.items.push_back(item());
          }
        }
        break;
      }
      case 6:  // a tag: the tagged item stands for itself here
        arg(ai, indef);
        FEEC_REQUIRE(!indef, FEEC_E_ARG, "malformed CBOR: indefinite tag");
        c = item();
        break;
      default:  // 7: floats and simple values
        if (ai == 25) {
          c.kind = Cbor::FLOAT;
          c.f = half_to_double((uint16_t)be(2));
        } else if (ai == 26) {
          const uint32_t b = (uint32_t)be(4);
          float f;
          std::memcpy(&f, &b, 4);
          c.kind = Cbor::FLOAT;
          c.f = f;
        } else if (ai == 27) {
          const uint64_t b = be(8);
          std::memcpy(&c.f, &b, 8);
          c.kind = Cbor::FLOAT;
        } else {
          FEEC_REQUIRE(ai != 31, FEEC_E_ARG, "malformed CBOR: unexpected break");
          c.kind = Cbor::SIMPLE;
          c.u = ai < 24 ? ai : arg(ai, indef);
        }
    }
    --depth;
    return c;
  }
};

Cbor parse_cbor(const uint8_t* bytes, int64_t len) {
  FEEC_REQUIRE((bytes || len == 0) && len >= 0, FEEC_E_ARG, "null bytes");
  CborReader r{bytes, bytes + len};
  Cbor c = r.item();
  FEEC_REQUIRE(r.p == r.end, FEEC_E_ARG, "malformed CBOR: bytes after the item");
  return c;
}

const Cbor* field(const Cbor& m, const char* key) {
  if (m.kind != Cbor::MAP) return nullptr;
  for (size_t i = 0; i + 1 < m.items.size(); i += 2)
    if (m.items[i].kind == Cbor::TEXT && m.items[i].s == key) return &m.items[i + 1];
  return nullptr;
}

double number(const Cbor& c) {
  if (c.kind == Cbor::FLOAT) return c.f;
  if (c.kind == Cbor::UINT) return (double)c.u;
  if (c.kind == Cbor::NINT) return -1.0 - (double)c.u;
  fail(FEEC_E_ARG, "CBOR: expected a number");
}

// cells of a stored skeleton: array of {"vertices": [...]} (or of plain arrays)
void cbor_cells(const Cbor& top, int& dim, std::vector<int64_t>& cells, int64_t& ncells) {
  FEEC_REQUIRE(top.kind == Cbor::ARRAY, FEEC_E_ARG, "CBOR complex: expected an array of simplices");
  FEEC_REQUIRE(!top.items.empty(), FEEC_E_ARG, "CBOR complex: no simplices");
  dim = -1;
  cells.clear();
  for (const Cbor& s : top.items) {
    const Cbor* v = s.kind == Cbor::ARRAY ? &s : field(s, "vertices");
    FEEC_REQUIRE(v && v->kind == Cbor::ARRAY && !v->items.empty(), FEEC_E_ARG,
                 "CBOR complex: a simplex without a vertex array");
    const int d = (int)v->items.size() - 1;
    FEEC_REQUIRE(dim < 0 || d == dim, FEEC_E_ARG, "CBOR complex: simplices of different dimensions");
    FEEC_REQUIRE(d <= 4, FEEC_E_UNSUPPORTED, "CBOR complex: dimension above 4");
    dim = d;
    for (const Cbor& x : v->items) {
      FEEC_REQUIRE(x.kind == Cbor::UINT && x.u < (uint64_t(1) << 62), FEEC_E_ARG,
                   "CBOR complex: a vertex is not an unsigned integer");
      cells.push_back((int64_t)x.u);
    }
  }
  ncells = (int64_t)top.items.size();
}

// the vector of a stored MeshLengthsSq
const Cbor& cbor_vector_data(const Cbor& top) {
  const Cbor* v = field(top, "vector");
  FEEC_REQUIRE(v, FEEC_E_ARG, "CBOR lengths: no \"vector\" field");
  if (v->kind == Cbor::ARRAY) {
    // a plain array of numbers, or the storage as a sequence (data, nrows, ncols)
    if (!v->items.empty() && v->items[0].kind == Cbor::ARRAY) return v->items[0];
    return *v;
  }
  const Cbor* data = field(*v, "data");
  FEEC_REQUIRE(data && data->kind == Cbor::ARRAY, FEEC_E_ARG, "CBOR lengths: no \"data\" array");
  const Cbor* nrows = field(*v, "nrows");
  if (nrows && nrows->kind == Cbor::UINT)
    FEEC_REQUIRE(nrows->u == data->items.size(), FEEC_E_ARG, "CBOR lengths: nrows != data length");
  return *data;
}

struct CborWriter {
  std::vector<uint8_t> out;
  void head(int major, uint64_t v) {
    const uint8_t m = (uint8_t)(major << 5);
    if (v < 24) {
      out.push_back(m | (uint8_t)v);
    } else if (v <= 0xff) {
      out.push_back(m | 24);
      out.push_back((uint8_t)v);
    } else if (v <= 0xffff) {
      out.push_back(m | 25);
      for (int i = 1; i >= 0; --i) out.push_back((uint8_t)(v >> (8 * i)));
    } else if (v <= 0xffffffffull) {
      out.push_back(m | 26);
      for (int i = 3; i >= 0; --i) out.push_back((uint8_t)(v >> (8 * i)));
    } else {
      out.push_back(m | 27);
      for (int i = 7; i >= 0; --i) out.push_back((uint8_t)(v >> (8 * i)));
    }
  }
  void text(const char* s) {
    const size_t n = std::strlen(s);
    head(3, n);
    out.insert(out.end(), s, s + n);
  }
  // the shortest of binary16 / binary32 / binary64 that holds v exactly
  void real(double v) {
    if (!std::isnan(v)) {
      const float f = (float)v;
      if ((double)f == v || std::isinf(v)) {
        uint16_t h;
        if (double_to_half_exact(v, h)) {
          out.push_back(0xf9);
          out.push_back((uint8_t)(h >> 8));
          out.push_back((uint8_t)h);
          return;
        }
        uint32_t b;
        std::memcpy(&b, &f, 4);
        out.push_back(0xfa);
        for (int i = 3; i >= 0; --i) out.push_back((uint8_t)(b >> (8 * i)));
        return;
      }
    }
    uint64_t b;
    std::memcpy(&b, &v, 8);
    out.push_back(0xfb);
    for (int i = 7; i >= 0; --i) out.push_back((uint8_t)(b >> (8 * i)));
  }
  // v as binary16 if that is exact (zeros, infinities, normals and subnormals)
  static bool double_to_half_exact(double v, uint16_t& h) {
    const uint16_t sign = std::signbit(v) ? 0x8000 : 0;
    if (v == 0.0) {
      h = sign;
      return true;
    }
    if (std::isinf(v)) {
      h = sign | 0x7c00;
      return true;
    }
    const double a = std::fabs(v);
    int e;
    const double m = std::frexp(a, &e);  // a = m 2^e, m in [0.5, 1)
    // normal halves: 2^-14 <= a < 65520 with 11 significant bits
    if (e - 1 >= -14 && e - 1 <= 15) {
      const double scaled = std::ldexp(m, 11);  // integer in [1024, 2048) if exact
      if (scaled != std::floor(scaled)) return false;
      h = sign | (uint16_t)((e - 1 + 15) << 10) | (uint16_t)((int)scaled - 1024);
      return true;
    }
    // subnormal halves: multiples of 2^-24 below 2^-14
    if (e - 1 < -14) {
      const double q = std::ldexp(a, 24);
      if (q != std::floor(q) || q >= 1024.0) return false;
      h = sign | (uint16_t)q;
      return true;
    }
    return false;
  }
};

}  // namespace
}  // namespace feec

using namespace feec;

extern "C" {

feec_status feec_gmsh_read(const char* bytes, int64_t len, int* dim, int* coord_dim,
                           int64_t* nverts, int64_t* ncells, int64_t* cells, double* coords) {
  return guard([&] {
    FEEC_REQUIRE(dim && coord_dim && nverts && ncells && len >= 0, FEEC_E_ARG, "null argument");
    Gmsh g = read_gmsh(bytes, len);
    *dim = g.dim;
    *coord_dim = g.coord_dim;
    *nverts = g.nverts;
    *ncells = g.ncells;
    if (cells) std::copy(g.words.begin(), g.words.end(), cells);
    if (coords) std::copy(g.coords.begin(), g.coords.end(), coords);
  });
}

feec_status feec_cbor_read_cells(const uint8_t* bytes, int64_t len, int* dim, int64_t* ncells,
                                 int64_t* cells) {
  return guard([&] {
    FEEC_REQUIRE(dim && ncells, FEEC_E_ARG, "null argument");
    std::vector<int64_t> c;
    int64_t n = 0;
    cbor_cells(parse_cbor(bytes, len), *dim, c, n);
    *ncells = n;
    if (cells) std::copy(c.begin(), c.end(), cells);
  });
}

feec_status feec_cbor_read_lengths(const uint8_t* bytes, int64_t len, int64_t* nedges,
                                   double* lengths_sq) {
  return guard([&] {
    FEEC_REQUIRE(nedges, FEEC_E_ARG, "null argument");
    const Cbor top = parse_cbor(bytes, len);
    const Cbor& data = cbor_vector_data(top);
    *nedges = (int64_t)data.items.size();
    if (lengths_sq)
      for (size_t i = 0; i < data.items.size(); ++i) lengths_sq[i] = number(data.items[i]);
  });
}

feec_status feec_cbor_write_cells(int dim, int64_t ncells, const int64_t* cells, uint8_t* out,
                                  int64_t* len) {
  return guard([&] {
    FEEC_REQUIRE(len && dim >= 0 && dim <= 4 && ncells > 0 && cells, FEEC_E_ARG,
                 "dimension 0..4, at least one cell");
    CborWriter w;
    w.head(4, (uint64_t)ncells);
    for (int64_t c = 0; c < ncells; ++c) {
      w.head(5, 1);
      w.text("vertices");
      w.head(4, (uint64_t)(dim + 1));
      for (int i = 0; i <= dim; ++i) {
        const int64_t v = cells[c * (dim + 1) + i];
        FEEC_REQUIRE(v >= 0, FEEC_E_ARG, "negative vertex id");
        w.head(0, (uint64_t)v);
      }
    }
    if (out) {
      FEEC_REQUIRE(*len >= (int64_t)w.out.size(), FEEC_E_ARG, "output buffer too small");
      std::copy(w.out.begin(), w.out.end(), out);
    }
    *len = (int64_t)w.out.size();
  });
}

feec_status feec_cbor_write_lengths(int64_t nedges, const double* lengths_sq, uint8_t* out,
                                    int64_t* len) {
  return guard([&] {
    FEEC_REQUIRE(len && nedges >= 0 && (nedges == 0 || lengths_sq), FEEC_E_ARG, "null argument");
    CborWriter w;
    w.head(5, 1);
    w.text("vector");
    w.head(5, 3);
    w.text("data");
    w.head(4, (uint64_t)nedges);
    for (int64_t i = 0; i < nedges; ++i) w.real(lengths_sq[i]);
    w.text("nrows");
    w.head(0, (uint64_t)nedges);
    w.text("ncols");
    w.out.push_back(0xf6);  // null: the static column count
    if (out) {
      FEEC_REQUIRE(*len >= (int64_t)w.out.size(), FEEC_E_ARG, "output buffer too small");
      std::copy(w.out.begin(), w.out.end(), out);
    }
    *len = (int64_t)w.out.size();
  });
}

}  // extern "C"
