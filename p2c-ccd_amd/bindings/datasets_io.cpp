// p2cccd_cpp: the on-disk inputs of the certificate path, read natively.
//
//   read_tight_inclusion_csv  Tight-Inclusion rational CSV in the rule of the
//       reference's C++ query benchmark (src/tools/
//       tight_inclusion_full_query_benchmark.cpp:291-321, 412-455): blocks of
//       8 rows (4 vertices at t0, 4 at t1), a row = x_num,x_den,y_num,y_den,
//       z_num,z_den,truth; each coordinate is the long-double quotient of the
//       two parsed integers, rounded to double; the truth column is "0" or
//       "1" and equal over a block; a trailing incomplete block is ignored.
//   load_binary_ply  binary PLY frames as the reference's scene tool reads
//       them (src/tools/scene_object_envelope_native_ti_walltime.cpp:
//       338-489): the header's vertex properties in order (x, y, z kept,
//       others read and dropped), the face list's count and index types,
//       polygons fanned from their first corner, triangles with a repeated
//       corner dropped.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cerrno>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace py = pybind11;

namespace {

// ---- Tight-Inclusion CSV ---------------------------------------------------

// The fields of a row split at ',' the way repeated std::getline(stream,
// token, ',') splits it: no field after a final ',' and none for an empty
// row.  Trailing '\r', ' ' and '\t' are cut from each field.
std::vector<std::string> CsvFields(const std::string& line) {
  std::vector<std::string> out;
  std::size_t start = 0;
  while (start < line.size()) {
    const std::size_t comma = line.find(',', start);
    std::string f = line.substr(start, comma == std::string::npos ? std::string::npos : comma - start);
    while (!f.empty() && (f.back() == '\r' || f.back() == ' ' || f.back() == '\t')) f.pop_back();
    out.push_back(std::move(f));
    if (comma == std::string::npos) break;
    start = comma + 1;
  }
  return out;
}

// std::stold's rule: strtold over the field, an error when nothing parses
// or the value is out of range.
long double ParseLongDouble(const std::string& s) {
  const char* begin = s.c_str();
  char* end = nullptr;
  errno = 0;
  const long double v = std::strtold(begin, &end);
  if (end == begin) throw std::invalid_argument("invalid Tight-Inclusion CSV number: " + s);
  if (errno == ERANGE) throw std::out_of_range("Tight-Inclusion CSV number out of range: " + s);
  return v;
}

void ParseRow(const std::string& line, double* xyz, bool* truth) {
  const std::vector<std::string> f = CsvFields(line);
  if (f.size() != 7) throw std::runtime_error("Tight-Inclusion CSV row must contain 7 columns");
  for (int k = 0; k < 3; ++k) {
    const long double q = ParseLongDouble(f[2 * k]) / ParseLongDouble(f[2 * k + 1]);
    xyz[k] = static_cast<double>(q);
  }
  if (f[6] != "0" && f[6] != "1") throw std::runtime_error("truth column must be 0 or 1");
  *truth = f[6] == "1";
}

py::tuple ReadTightInclusionCsv(const std::string& path, std::uint64_t max_queries) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("failed to open CSV " + path);
  std::vector<double> xs;
  std::vector<std::uint8_t> truths;
  std::string rows[8];
  for (;;) {
    if (max_queries != 0 && truths.size() >= max_queries) break;
    bool block = true;
    for (int r = 0; r < 8 && block; ++r) block = static_cast<bool>(std::getline(in, rows[r]));
    if (!block) break;
    double q[24];
    bool first = false;
    for (int r = 0; r < 8; ++r) {
      bool t = false;
      ParseRow(rows[r], q + 3 * r, &t);
      if (r == 0) first = t;
      else if (t != first) throw std::runtime_error("truth label changes inside query block in " + path);
    }
    xs.insert(xs.end(), q, q + 24);
    truths.push_back(first ? 1 : 0);
  }
  const py::ssize_t n = static_cast<py::ssize_t>(truths.size());
  py::array_t<double> x({n, py::ssize_t(8), py::ssize_t(3)});
  py::array_t<bool> t(n);
  if (n) {
    std::memcpy(x.mutable_data(), xs.data(), xs.size() * sizeof(double));
    bool* tp = t.mutable_data();
    for (py::ssize_t i = 0; i < n; ++i) tp[i] = truths[static_cast<std::size_t>(i)] != 0;
  }
  return py::make_tuple(x, t);
}

// ---- binary PLY -------------------------------------------------------------

std::string Trimmed(const std::string& s) {
  const char* ws = " \t\r\n";
  const std::size_t b = s.find_first_not_of(ws);
  if (b == std::string::npos) return "";
  const std::size_t e = s.find_last_not_of(ws);
  return s.substr(b, e - b + 1);
}

std::vector<std::string> Words(const std::string& s) {
  std::istringstream is(s);
  std::vector<std::string> w;
  for (std::string x; is >> x;) w.push_back(x);
  return w;
}

struct PlyIn {
  std::ifstream in;
  bool big = false;

  template <typename T>
  T Read() {
    unsigned char b[sizeof(T)];
    in.read(reinterpret_cast<char*>(b), sizeof(T));
    if (!in) throw std::runtime_error("unexpected EOF while reading binary PLY");
    if (big) {
      for (std::size_t i = 0; i < sizeof(T) / 2; ++i) std::swap(b[i], b[sizeof(T) - 1 - i]);
    }
    T v;
    std::memcpy(&v, b, sizeof(T));
    return v;
  }

  double Scalar(const std::string& type) {
    if (type == "double" || type == "float64") return Read<double>();
    if (type == "float" || type == "float32") return static_cast<double>(Read<float>());
    if (type == "int" || type == "int32") return static_cast<double>(Read<std::int32_t>());
    if (type == "uint" || type == "uint32") return static_cast<double>(Read<std::uint32_t>());
    if (type == "short" || type == "int16") return static_cast<double>(Read<std::int16_t>());
    if (type == "ushort" || type == "uint16") return static_cast<double>(Read<std::uint16_t>());
    if (type == "uchar" || type == "uint8") return static_cast<double>(Read<std::uint8_t>());
    if (type == "char" || type == "int8") return static_cast<double>(Read<std::int8_t>());
    throw std::runtime_error("unsupported PLY scalar type: " + type);
  }

  // list counts and indices: unsigned 32-bit, signed types sign-extended
  std::uint32_t Index(const std::string& type) {
    if (type == "uchar" || type == "uint8") return Read<std::uint8_t>();
    if (type == "char" || type == "int8") return static_cast<std::uint32_t>(Read<std::int8_t>());
    if (type == "ushort" || type == "uint16") return Read<std::uint16_t>();
    if (type == "short" || type == "int16") return static_cast<std::uint32_t>(Read<std::int16_t>());
    if (type == "uint" || type == "uint32") return Read<std::uint32_t>();
    if (type == "int" || type == "int32") return static_cast<std::uint32_t>(Read<std::int32_t>());
    throw std::runtime_error("unsupported PLY integer type: " + type);
  }
};

py::tuple LoadBinaryPly(const std::string& path) {
  PlyIn p;
  p.in.open(path, std::ios::binary);
  if (!p.in) throw std::runtime_error("failed to open PLY " + path);
  std::string line;
  std::getline(p.in, line);
  if (Trimmed(line) != "ply") throw std::runtime_error("not a PLY file: " + path);
  bool little = false;
  std::size_t n_vertices = 0, n_faces = 0;
  std::string element;
  std::vector<std::pair<std::string, std::string>> props;  // (name, type) of the vertex element
  std::string count_type = "uint8", index_type = "int32";
  while (std::getline(p.in, line)) {
    line = Trimmed(line);
    if (line == "end_header") break;
    const std::vector<std::string> w = Words(line);
    if (w.empty()) continue;
    if (w[0] == "format") {
      little = w.size() >= 2 && w[1] == "binary_little_endian";
      p.big = w.size() >= 2 && w[1] == "binary_big_endian";
    } else if (w[0] == "element" && w.size() >= 3) {
      element = w[1];
      if (element == "vertex") n_vertices = static_cast<std::size_t>(std::stoull(w[2]));
      else if (element == "face") n_faces = static_cast<std::size_t>(std::stoull(w[2]));
    } else if (w[0] == "property" && element == "vertex" && w.size() >= 3) {
      props.emplace_back(w[2], w[1]);
    } else if (w[0] == "property" && element == "face" && w.size() >= 5 && w[1] == "list") {
      count_type = w[2];
      index_type = w[3];
    }
  }
  if (!little && !p.big) throw std::runtime_error("only binary PLY is supported: " + path);
  if (n_vertices == 0 || n_faces == 0) throw std::runtime_error("PLY has no vertices or faces: " + path);
  std::vector<double> v(3 * n_vertices, 0.0);
  for (std::size_t i = 0; i < n_vertices; ++i) {
    for (const auto& [name, type] : props) {
      const double x = p.Scalar(type);
      if (name == "x") v[3 * i] = x;
      else if (name == "y") v[3 * i + 1] = x;
      else if (name == "z") v[3 * i + 2] = x;
    }
  }
  std::vector<std::int32_t> f;
  f.reserve(3 * n_faces);
  std::vector<int> corner;
  for (std::size_t i = 0; i < n_faces; ++i) {
    const std::uint32_t count = p.Index(count_type);
    corner.assign(count, 0);
    for (std::uint32_t k = 0; k < count; ++k) corner[k] = static_cast<int>(p.Index(index_type));
    for (std::uint32_t k = 1; k + 1 < count; ++k) {
      const int a = corner[0], b = corner[k], c = corner[k + 1];
      if (a != b && b != c && a != c) {
        f.push_back(a);
        f.push_back(b);
        f.push_back(c);
      }
    }
  }
  const py::ssize_t nv = static_cast<py::ssize_t>(n_vertices), nf = static_cast<py::ssize_t>(f.size() / 3);
  py::array_t<double> verts({nv, py::ssize_t(3)});
  py::array_t<std::int32_t> faces({nf, py::ssize_t(3)});
  std::memcpy(verts.mutable_data(), v.data(), v.size() * sizeof(double));
  if (nf) std::memcpy(faces.mutable_data(), f.data(), f.size() * sizeof(std::int32_t));
  return py::make_tuple(verts, faces);
}

}  // namespace

void BindDatasets(py::module_& m) {
  m.def("read_tight_inclusion_csv", &ReadTightInclusionCsv, py::arg("path"), py::arg("max_queries") = 0,
        "Queries of a Tight-Inclusion rational CSV, parsed as the reference's C++ query benchmark parses them: "
        "(vertices float64[n, 8, 3], ground truth bool[n]).");
  m.def("load_binary_ply", &LoadBinaryPly, py::arg("path"),
        "A binary PLY frame as the reference's scene tool reads it: (vertices float64[V, 3], "
        "triangles int32[F, 3]).");
}
