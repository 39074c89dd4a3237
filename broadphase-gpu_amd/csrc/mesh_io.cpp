// mesh_io.cpp -- host-side mesh input and ground-truth comparison.
//
// Replaces the reference's parseMesh (src/io.cu:30-51: igl::readOBJ /
// igl::readPLY of the t0 and t1 meshes, then igl::edges) and
// compare_mathematica (src/groundtruth.cu:32-67: the JSON list of true
// pairs, reported as "Contains k/n TP").  libigl and nlohmann_json are not
// used; the formats are read directly:
//   OBJ : "v x y z" lines and "f a b c" triangles (1-based, negative =
//         relative, "a/t/n" forms accepted; other polygons are rejected)
//   PLY : ascii or binary_little_endian; vertex x, y, z of any numeric type,
//         face "vertex_indices" (or "vertex_index") lists of 3
//   JSON: an array of two-integer arrays
// Edges: every undirected edge of the faces once, (smaller, larger), ordered
// by larger index then smaller index -- the order in which igl::edges reads
// them off the column-major adjacency matrix.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <new>
#include <set>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "../../include/stq.h"

namespace {

struct MeshError {
  int status;
  std::string what;
};

[[noreturn]] void fail(const std::string &w, int status = STQ_ERR_INVALID_ARGUMENT) {
  throw MeshError{status, w};
}

bool read_obj(const std::string &path, std::vector<double> &V, std::vector<int32_t> &F) {
  std::ifstream in(path);
  if (!in) return false;
  std::string line;
  int64_t lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    std::istringstream ss(line);
    std::string tag;
    if (!(ss >> tag)) continue;
    if (tag == "v") {
      double x, y, z;
      if (!(ss >> x >> y >> z)) fail("OBJ line " + std::to_string(lineno) + ": bad vertex");
      V.push_back(x);
      V.push_back(y);
      V.push_back(z);
    } else if (tag == "f") {
      std::vector<int64_t> idx;
      std::string tok;
      while (ss >> tok) {
        const std::string head = tok.substr(0, tok.find('/'));
        char *end = nullptr;
        errno = 0;
        const long long v = std::strtoll(head.c_str(), &end, 10);
        if (head.empty() || *end != '\0' || errno != 0 || v == 0)
          fail("OBJ line " + std::to_string(lineno) + ": bad face index");
        const int64_t nv = (int64_t)(V.size() / 3);
        const int64_t i = v > 0 ? (int64_t)v - 1 : nv + (int64_t)v;
        // narrowed to int32 only in range; the vertex count is checked at the end
        if (i < 0 || i > INT32_MAX) fail("OBJ line " + std::to_string(lineno) + ": face index out of range");
        idx.push_back(i);
      }
      if (idx.size() != 3) fail("OBJ line " + std::to_string(lineno) + ": only triangles are read");
      for (int64_t i : idx) F.push_back((int32_t)i);
    }
  }
  return true;
}

// ---- PLY ----
enum PType { P_I8, P_U8, P_I16, P_U16, P_I32, P_U32, P_F32, P_F64, P_BAD };
PType ptype(const std::string &t) {
  if (t == "char" || t == "int8") return P_I8;
  if (t == "uchar" || t == "uint8") return P_U8;
  if (t == "short" || t == "int16") return P_I16;
  if (t == "ushort" || t == "uint16") return P_U16;
  if (t == "int" || t == "int32") return P_I32;
  if (t == "uint" || t == "uint32") return P_U32;
  if (t == "float" || t == "float32") return P_F32;
  if (t == "double" || t == "float64") return P_F64;
  return P_BAD;
}
int psize(PType t) {
  static const int s[] = {1, 1, 2, 2, 4, 4, 4, 8, 0};
  return s[t];
}
double pread_bin(const unsigned char *p, PType t) {
  switch (t) {
    case P_I8: { int8_t v; std::memcpy(&v, p, 1); return v; }
    case P_U8: return p[0];
    case P_I16: { int16_t v; std::memcpy(&v, p, 2); return v; }
    case P_U16: { uint16_t v; std::memcpy(&v, p, 2); return v; }
    case P_I32: { int32_t v; std::memcpy(&v, p, 4); return v; }
    case P_U32: { uint32_t v; std::memcpy(&v, p, 4); return v; }
    case P_F32: { float v; std::memcpy(&v, p, 4); return v; }
    case P_F64: { double v; std::memcpy(&v, p, 8); return v; }
    default: return 0;
  }
}

struct PProp {
  std::string name;
  bool list = false;
  PType count = P_BAD, type = P_BAD;
};
struct PElem {
  std::string name;
  int64_t n = 0;
  std::vector<PProp> props;
};

bool read_ply(const std::string &path, std::vector<double> &V, std::vector<int32_t> &F) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return false;
  std::string line;
  if (!std::getline(in, line) || line.substr(0, 3) != "ply") fail("not a PLY file");
  bool ascii = false;
  std::vector<PElem> elems;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::istringstream ss(line);
    std::string w;
    ss >> w;
    if (w == "format") {
      std::string f;
      ss >> f;
      if (f == "ascii") ascii = true;
      else if (f != "binary_little_endian") fail("PLY format " + f + " is not read");
    } else if (w == "element") {
      PElem e;
      ss >> e.name >> e.n;
      if (e.n < 0) fail("PLY: negative element count");
      elems.push_back(e);
    } else if (w == "property") {
      if (elems.empty()) fail("PLY: property before element");
      PProp p;
      std::string t;
      ss >> t;
      if (t == "list") {
        std::string c, v;
        ss >> c >> v >> p.name;
        p.list = true;
        p.count = ptype(c);
        p.type = ptype(v);
      } else {
        p.type = ptype(t);
        ss >> p.name;
      }
      if (p.type == P_BAD || (p.list && p.count == P_BAD)) fail("PLY: unknown property type");
      elems.back().props.push_back(p);
    } else if (w == "end_header") {
      break;
    }
  }
  std::vector<char> rest((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  size_t at = 0;
  std::istringstream asc(ascii ? std::string(rest.begin(), rest.end()) : std::string());
  auto next_value = [&](PType t) -> double {
    if (ascii) {
      double v;
      if (!(asc >> v)) fail("PLY: truncated ascii data");
      return v;
    }
    const int s = psize(t);
    if (at + s > rest.size()) fail("PLY: truncated binary data");
    const double v = pread_bin(reinterpret_cast<const unsigned char *>(rest.data()) + at, t);
    at += s;
    return v;
  };
  for (const PElem &e : elems) {
    int ix = -1, iy = -1, iz = -1, iface = -1;
    for (size_t k = 0; k < e.props.size(); ++k) {
      const std::string &nm = e.props[k].name;
      if (nm == "x") ix = (int)k;
      if (nm == "y") iy = (int)k;
      if (nm == "z") iz = (int)k;
      if ((nm == "vertex_indices" || nm == "vertex_index") && e.props[k].list) iface = (int)k;
    }
    const bool is_vertex = e.name == "vertex", is_face = e.name == "face";
    if (is_vertex && (ix < 0 || iy < 0 || iz < 0)) fail("PLY: vertex without x, y, z");
    if (is_face && iface < 0) fail("PLY: face without vertex_indices");
    for (int64_t r = 0; r < e.n; ++r) {
      double xyz[3] = {0, 0, 0};
      for (size_t k = 0; k < e.props.size(); ++k) {
        const PProp &p = e.props[k];
        if (p.list) {
          const double c = next_value(p.count);
          if (c < 0 || c > 1e6) fail("PLY: bad list length");
          const int64_t m = (int64_t)c;
          if ((int)k == iface && is_face && m != 3) fail("PLY: only triangles are read");
          for (int64_t j = 0; j < m; ++j) {
            const double v = next_value(p.type);
            if ((int)k == iface && is_face) {
              // an integral value in int32 range, checked before the conversion
              if (!(v >= 0.0 && v <= (double)INT32_MAX) || v != std::floor(v))
                fail("PLY: face index out of range");
              F.push_back((int32_t)v);
            }
          }
        } else {
          const double v = next_value(p.type);
          if ((int)k == ix) xyz[0] = v;
          if ((int)k == iy) xyz[1] = v;
          if ((int)k == iz) xyz[2] = v;
        }
      }
      if (is_vertex) V.insert(V.end(), xyz, xyz + 3);
    }
  }
  return true;
}

template <class T>
T *to_malloc(const std::vector<T> &v) {
  T *p = static_cast<T *>(std::malloc(sizeof(T) * std::max<size_t>(v.size(), 1)));
  if (!p) throw std::bad_alloc();
  if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
  return p;
}

template <class Fn>
int guard(Fn &&fn) {
  try {
    fn();
    return STQ_OK;
  } catch (const MeshError &e) {
    std::fprintf(stderr, "stq: %s\n", e.what.c_str());
    return e.status;
  } catch (const std::bad_alloc &) {
    return STQ_ERR_OUT_OF_MEMORY;
  } catch (...) {
    return STQ_ERR_INTERNAL;
  }
}

// [[a, b], [c, d], ...] -> pairs
std::vector<std::pair<int64_t, int64_t>> read_pair_json(const std::string &path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) fail("cannot open " + path);
  const std::string s((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  size_t i = 0;
  auto ws = [&]() {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
  };
  auto expect = [&](char c) {
    ws();
    if (i >= s.size() || s[i] != c) fail(std::string("JSON: expected '") + c + "'");
    ++i;
  };
  auto integer = [&]() -> int64_t {
    ws();
    size_t j = i;
    if (j < s.size() && (s[j] == '-' || s[j] == '+')) ++j;
    const size_t d0 = j;
    while (j < s.size() && std::isdigit((unsigned char)s[j])) ++j;
    if (j == d0) fail("JSON: expected an integer");
    const std::string tok = s.substr(i, j - i);
    errno = 0;
    char *end = nullptr;
    const long long v = std::strtoll(tok.c_str(), &end, 10);
    if (errno != 0 || *end != '\0') fail("JSON: integer out of range");
    i = j;
    return (int64_t)v;
  };
  std::vector<std::pair<int64_t, int64_t>> out;
  expect('[');
  ws();
  if (i < s.size() && s[i] == ']') return out;
  while (true) {
    expect('[');
    const int64_t a = integer();
    expect(',');
    const int64_t b = integer();
    expect(']');
    out.emplace_back(a, b);
    ws();
    if (i < s.size() && s[i] == ',') {
      ++i;
      continue;
    }
    expect(']');
    break;
  }
  return out;
}

}  // namespace

extern "C" {

int stq_mesh_read(const char *path, double **h_vertices, int64_t *nv, int32_t **h_faces,
                  int64_t *nf) {
  if (!path || !h_vertices || !nv || !h_faces || !nf) return STQ_ERR_INVALID_ARGUMENT;
  *h_vertices = nullptr;
  *h_faces = nullptr;
  *nv = *nf = 0;
  return guard([&] {
    std::string p(path);
    const size_t dot = p.rfind('.');
    std::string ext = dot == std::string::npos ? "" : p.substr(dot + 1);
    for (char &c : ext) c = (char)std::tolower((unsigned char)c);
    std::vector<double> V;
    std::vector<int32_t> F;
    bool ok = false;
    if (ext == "obj") ok = read_obj(p, V, F);
    else if (ext == "ply") ok = read_ply(p, V, F);
    else fail("unknown mesh format (.obj or .ply)");
    if (!ok) fail("cannot open " + p);
    const int64_t n = (int64_t)(V.size() / 3);
    for (int32_t f : F)
      if (f < 0 || f >= n) fail("face index outside the vertex list");
    *h_vertices = to_malloc(V);
    *h_faces = to_malloc(F);
    *nv = n;
    *nf = (int64_t)(F.size() / 3);
  });
}

int stq_mesh_edges(const int32_t *h_faces, int64_t nf, int32_t **h_edges, int64_t *ne) {
  if ((nf > 0 && !h_faces) || nf < 0 || !h_edges || !ne) return STQ_ERR_INVALID_ARGUMENT;
  *h_edges = nullptr;
  *ne = 0;
  return guard([&] {
    std::vector<std::pair<int32_t, int32_t>> e;  // (larger, smaller)
    e.reserve((size_t)nf * 3);
    for (int64_t i = 0; i < nf; ++i)
      for (int k = 0; k < 3; ++k) {
        const int32_t a = h_faces[3 * i + k], b = h_faces[3 * i + (k + 1) % 3];
        if (a < 0 || b < 0) fail("negative face index");
        if (a != b) e.emplace_back(std::max(a, b), std::min(a, b));
      }
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    std::vector<int32_t> out;
    out.reserve(e.size() * 2);
    for (const auto &x : e) {
      out.push_back(x.second);
      out.push_back(x.first);
    }
    *h_edges = to_malloc(out);
    *ne = (int64_t)e.size();
  });
}

int stq_recall(const stq_pair *h_pairs, int64_t count, const char *json_path, int32_t canonical,
               int64_t *true_positives, int64_t *total) {
  if ((count > 0 && !h_pairs) || count < 0 || !json_path || !true_positives || !total)
    return STQ_ERR_INVALID_ARGUMENT;
  return guard([&] {
    auto truth = read_pair_json(json_path);
    auto key = [&](int64_t a, int64_t b) {
      return canonical ? std::make_pair(std::min(a, b), std::max(a, b)) : std::make_pair(a, b);
    };
    std::set<std::pair<int64_t, int64_t>> t;
    for (const auto &p : truth) t.insert(key(p.first, p.second));
    std::set<std::pair<int64_t, int64_t>> mine;
    for (int64_t k = 0; k < count; ++k) mine.insert(key(h_pairs[k].first, h_pairs[k].second));
    int64_t tp = 0;
    for (const auto &p : t) tp += mine.count(p) ? 1 : 0;
    *true_positives = tp;
    *total = (int64_t)t.size();
  });
}

}  // extern "C"
