// ply_reader.cc -- PLY ingestion (include/phaser_ply.h): the header is
// parsed into elements and typed properties, the body is read element by
// element (ASCII tokens or packed binary of either byte order), and the
// vertex element's requested fields are converted to double and then to the
// cloud's float, as the reference's tinyply request into std::vector<double>
// followed by parsePlyPointCloud does (ply-helper.cc:25-65,
// point-cloud.cc:330-349).
#include "phaser_ply.h"

#include <cstdio>
#include <cstdlib>
#include <exception>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace {

thread_local std::string g_ply_error;

int fail(const std::string& msg) {
  g_ply_error = msg;
  return -1;
}

enum class Type { kI8, kU8, kI16, kU16, kI32, kU32, kF32, kF64, kBad };

Type parse_type(const std::string& t) {
  if (t == "char" || t == "int8") return Type::kI8;
  if (t == "uchar" || t == "uint8") return Type::kU8;
  if (t == "short" || t == "int16") return Type::kI16;
  if (t == "ushort" || t == "uint16") return Type::kU16;
  if (t == "int" || t == "int32") return Type::kI32;
  if (t == "uint" || t == "uint32") return Type::kU32;
  if (t == "float" || t == "float32") return Type::kF32;
  if (t == "double" || t == "float64") return Type::kF64;
  return Type::kBad;
}

int type_size(Type t) {
  switch (t) {
    case Type::kI8: case Type::kU8: return 1;
    case Type::kI16: case Type::kU16: return 2;
    case Type::kI32: case Type::kU32: case Type::kF32: return 4;
    case Type::kF64: return 8;
    default: return 0;
  }
}

struct Property {
  std::string name;
  Type type = Type::kBad;
  bool list = false;
  Type count_type = Type::kBad;
};

struct Element {
  std::string name;
  int64_t count = 0;
  std::vector<Property> props;
};

enum class Format { kAscii, kBinaryLE, kBinaryBE };

// one binary value of type t at p (byte order: swap when big-endian on
// this little-endian host), as double
double read_binary(const unsigned char* p, Type t, bool swap) {
  unsigned char b[8];
  const int n = type_size(t);
  for (int i = 0; i < n; ++i) b[i] = swap ? p[n - 1 - i] : p[i];
  switch (t) {
    case Type::kI8: { int8_t v; std::memcpy(&v, b, 1); return v; }
    case Type::kU8: { uint8_t v; std::memcpy(&v, b, 1); return v; }
    case Type::kI16: { int16_t v; std::memcpy(&v, b, 2); return v; }
    case Type::kU16: { uint16_t v; std::memcpy(&v, b, 2); return v; }
    case Type::kI32: { int32_t v; std::memcpy(&v, b, 4); return v; }
    case Type::kU32: { uint32_t v; std::memcpy(&v, b, 4); return v; }
    case Type::kF32: { float v; std::memcpy(&v, b, 4); return v; }
    case Type::kF64: { double v; std::memcpy(&v, b, 8); return v; }
    default: return 0.0;
  }
}

}  // namespace

extern "C" {

void phz_ply_default_fields(phz_ply_fields* f) {
  if (!f) return;
  f->intensity = "intensity";
  f->reflectivity = "reflectivity";
  f->ambient = "noise";
  f->range = "range";
}

const char* phz_ply_last_error(void) { return g_ply_error.c_str(); }

void phz_ply_free(phz_ply_cloud* c) {
  if (!c) return;
  std::free(c->xyzi);
  std::free(c->extra);
  std::free(c->range);
  std::memset(c, 0, sizeof(*c));
}

static int ply_read(const char* path, const phz_ply_fields* fields, phz_ply_cloud* out);

int phz_ply_read(const char* path, const phz_ply_fields* fields, phz_ply_cloud* out) {
  // no C++ exception may cross the C ABI (a header's vertex count can ask
  // for more memory than the host has)
  try {
    return ply_read(path, fields, out);
  } catch (const std::exception& e) {
    if (out) phz_ply_free(out);
    return fail(std::string("reading the PLY file failed: ") + e.what());
  }
}

static int ply_read(const char* path, const phz_ply_fields* fields, phz_ply_cloud* out) {
  if (!path || !out) return fail("null argument");
  std::memset(out, 0, sizeof(*out));
  phz_ply_fields names;
  phz_ply_default_fields(&names);
  if (fields) {
    if (fields->intensity) names.intensity = fields->intensity;
    if (fields->reflectivity) names.reflectivity = fields->reflectivity;
    if (fields->ambient) names.ambient = fields->ambient;
    if (fields->range) names.range = fields->range;
  }
  std::ifstream in(path, std::ios::binary);
  if (!in) return fail(std::string("unable to open ply file: ") + path);
  std::vector<char> data((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());

  // header
  size_t pos = 0;
  auto next_line = [&](std::string* line) {
    if (pos >= data.size()) return false;
    size_t e = pos;
    while (e < data.size() && data[e] != '\n') ++e;
    line->assign(data.data() + pos, e - pos);
    if (!line->empty() && line->back() == '\r') line->pop_back();
    pos = e + 1;
    return true;
  };
  std::string line;
  if (!next_line(&line) || line != "ply") return fail("not a PLY file (no 'ply' magic)");
  Format format = Format::kAscii;
  bool have_format = false, ended = false;
  std::vector<Element> elements;
  while (next_line(&line)) {
    std::istringstream ls(line);
    std::string key;
    ls >> key;
    if (key == "format") {
      std::string f;
      ls >> f;
      if (f == "ascii") format = Format::kAscii;
      else if (f == "binary_little_endian") format = Format::kBinaryLE;
      else if (f == "binary_big_endian") format = Format::kBinaryBE;
      else return fail("unknown PLY format: " + f);
      have_format = true;
    } else if (key == "element") {
      Element e;
      ls >> e.name >> e.count;
      if (!ls || e.count < 0) return fail("malformed element line: " + line);
      elements.push_back(e);
    } else if (key == "property") {
      if (elements.empty()) return fail("property before any element");
      Property p;
      std::string t;
      ls >> t;
      if (t == "list") {
        std::string ct, it;
        ls >> ct >> it >> p.name;
        p.list = true;
        p.count_type = parse_type(ct);
        p.type = parse_type(it);
        if (p.count_type == Type::kBad || p.count_type == Type::kF32 ||
            p.count_type == Type::kF64)
          return fail("bad list count type: " + line);
      } else {
        p.type = parse_type(t);
        ls >> p.name;
      }
      if (p.type == Type::kBad || p.name.empty()) return fail("bad property line: " + line);
      elements.back().props.push_back(p);
    } else if (key == "end_header") {
      ended = true;
      break;
    } else if (key == "comment" || key == "obj_info" || key.empty()) {
      continue;
    } else {
      return fail("unknown header line: " + line);
    }
  }
  if (!have_format || !ended) return fail("incomplete PLY header");

  // the vertex element and the requested properties
  int vi = -1;
  for (size_t i = 0; i < elements.size(); ++i)
    if (elements[i].name == "vertex") vi = (int)i;
  if (vi < 0) return fail("no vertex element");
  const Element& V = elements[vi];
  auto find = [&](const char* name) {
    for (size_t i = 0; i < V.props.size(); ++i)
      if (!V.props[i].list && V.props[i].name == name) return (int)i;
    return -1;
  };
  const int px = find("x"), py = find("y"), pz = find("z");
  const int pi = find(names.intensity), pr = find(names.reflectivity);
  const int pa = find(names.ambient), pg = find(names.range);
  // ply-helper.cc:34-40: CHECK_GT on the xyz and intensity counts
  if (px < 0 || py < 0 || pz < 0 || V.count == 0) return fail("no x/y/z vertex data");
  if (pi < 0) return fail(std::string("no intensity field '") + names.intensity + "'");

  const int64_t n = V.count;
  // a binary body holds at least one byte per vertex property
  if (format != Format::kAscii && n > (int64_t)(data.size() - pos))
    return fail("vertex count exceeds the file size");
  std::vector<double> row(V.props.size());
  std::vector<float> xyzi((size_t)n * 4), extra((size_t)n * 2, 0.f);
  std::vector<double> range(pg >= 0 ? (size_t)n : 0);
  auto take_vertex = [&](int64_t k) {
    xyzi[4 * k + 0] = (float)row[px];
    xyzi[4 * k + 1] = (float)row[py];
    xyzi[4 * k + 2] = (float)row[pz];
    xyzi[4 * k + 3] = (float)row[pi];
    if (pr >= 0) extra[2 * k + 0] = (float)row[pr];
    if (pa >= 0) extra[2 * k + 1] = (float)row[pa];
    if (pg >= 0) range[k] = row[pg];
  };

  if (format == Format::kAscii) {
    // whitespace-separated values, element by element
    const char* p = data.data() + pos;
    const char* end = data.data() + data.size();
    // a float property is parsed as float, as tinyply reads it into its
    // declared type before the conversion to double
    auto token = [&](double* v, Type t) {
      while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
      if (p >= end) return false;
      char* q = nullptr;
      *v = t == Type::kF32 ? (double)std::strtof(p, &q) : std::strtod(p, &q);
      if (q == p) return false;
      p = q;
      return true;
    };
    for (size_t ei = 0; ei < elements.size(); ++ei) {
      const Element& E = elements[ei];
      for (int64_t k = 0; k < E.count; ++k)
        for (size_t j = 0; j < E.props.size(); ++j) {
          double v;
          const Property& P = E.props[j];
          if (!token(&v, P.list ? P.count_type : P.type))
            return fail("truncated ASCII body in element " + E.name);
          if (P.list) {
            const int64_t m = (int64_t)v;
            for (int64_t t = 0; t < m; ++t)
              if (!token(&v, P.type)) return fail("truncated list in element " + E.name);
          } else if ((int)ei == vi) {
            row[j] = v;
          }
          if ((int)ei == vi && j + 1 == E.props.size()) take_vertex(k);
        }
    }
  } else {
    const bool swap = format == Format::kBinaryBE;
    const unsigned char* p = reinterpret_cast<const unsigned char*>(data.data()) + pos;
    const unsigned char* end = reinterpret_cast<const unsigned char*>(data.data()) + data.size();
    for (size_t ei = 0; ei < elements.size(); ++ei) {
      const Element& E = elements[ei];
      for (int64_t k = 0; k < E.count; ++k) {
        for (size_t j = 0; j < E.props.size(); ++j) {
          const Property& P = E.props[j];
          if (P.list) {
            const int cs = type_size(P.count_type);
            if (end - p < cs) return fail("truncated binary body in element " + E.name);
            const int64_t m = (int64_t)read_binary(p, P.count_type, swap);
            p += cs;
            const int64_t skip = m * type_size(P.type);
            if (m < 0 || end - p < skip) return fail("truncated list in element " + E.name);
            p += skip;
          } else {
            const int s = type_size(P.type);
            if (end - p < s) return fail("truncated binary body in element " + E.name);
            if ((int)ei == vi) row[j] = read_binary(p, P.type, swap);
            p += s;
          }
        }
        if ((int)ei == vi) take_vertex(k);
      }
    }
  }

  out->n = n;
  out->xyzi = (float*)std::malloc(sizeof(float) * xyzi.size());
  out->extra = (float*)std::malloc(sizeof(float) * extra.size());
  if (!out->xyzi || !out->extra) {
    phz_ply_free(out);
    return fail("out of host memory");
  }
  std::memcpy(out->xyzi, xyzi.data(), sizeof(float) * xyzi.size());
  std::memcpy(out->extra, extra.data(), sizeof(float) * extra.size());
  if (pg >= 0) {
    out->range = (double*)std::malloc(sizeof(double) * range.size());
    if (!out->range) {
      phz_ply_free(out);
      return fail("out of host memory");
    }
    std::memcpy(out->range, range.data(), sizeof(double) * range.size());
  }
  out->has_reflectivity = pr >= 0;
  out->has_ambient = pa >= 0;
  out->has_range = pg >= 0;
  return 0;
}

int phz_ply_write_xyzi(const char* path, int64_t n, const float* xyzi) {
  if (!path || n < 0 || (n > 0 && !xyzi)) return fail("null argument");
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(std::string("unable to write ply file: ") + path);
  std::fprintf(f,
               "ply\nformat ascii 1.0\nelement vertex %lld\nproperty float x\nproperty float y\n"
               "property float z\nproperty float intensity\nend_header\n",
               (long long)n);
  for (int64_t i = 0; i < n; ++i)
    std::fprintf(f, "%.9g %.9g %.9g %.9g\n", xyzi[4 * i], xyzi[4 * i + 1], xyzi[4 * i + 2],
                 xyzi[4 * i + 3]);
  const bool ok = std::fclose(f) == 0;
  return ok ? 0 : fail(std::string("writing ply file failed: ") + path);
}

}  // extern "C"
