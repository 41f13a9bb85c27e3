// Minimal VTK XML PolyData (.vtp) reading and writing for oriented-disc point
// clouds, without VTK.
//
// Reading follows what the reference's vtp_point_cloud_reader extracts
// (src/rti/io/vtp_point_cloud_reader.hpp:23-95): the points, the cell-data
// normals (normalised when their length is off by more than 1e-6,
// util/utils.hpp:77-82), and the disc radius from the cell-data array
// "radius", else from "sqrt-of-area" times sqrt(3)/2 * (1 + 1/32). Data arrays
// may be ascii, inline binary or appended (base64 or raw), with UInt32 or
// UInt64 headers, uncompressed or zlib-compressed (vtkZLibDataCompressor) --
// the forms vtkXMLPolyDataWriter produces.
//
// Writing produces an ascii file with the reference writer's layout
// (src/rti/io/vtp_writer.hpp:26-41, 111-237, 303-335, 339-364): one vertex
// cell per disc, cell data Normals, radius, value, exposed-area, hitcnt,
// relative-error, variance-of-variance, and string field data; or a triangle
// surface with cell normals (the boundary box).
#pragma once

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace rti_vtk {

struct xml_tag {
  std::string name;
  std::map<std::string, std::string> attr;
  bool closing = false, self_closing = false;
  size_t begin = 0, end = 0;  // [begin, end) of the tag text in the file
};

// Next tag at or after pos (skipping comments and declarations); false at the end.
inline bool next_tag(const std::string& s, size_t& pos, xml_tag& t, size_t limit) {
  for (;;) {
    size_t a = s.find('<', pos);
    if (a == std::string::npos || a >= limit) return false;
    if (s.compare(a, 4, "<!--") == 0) {
      size_t e = s.find("-->", a);
      if (e == std::string::npos) return false;
      pos = e + 3;
      continue;
    }
    if (s.compare(a, 2, "<?") == 0 || s.compare(a, 2, "<!") == 0) {
      size_t e = s.find('>', a);
      if (e == std::string::npos) return false;
      pos = e + 1;
      continue;
    }
    size_t e = a + 1;
    char q = 0;
    while (e < s.size() && (q || s[e] != '>')) {
      if (q) {
        if (s[e] == q) q = 0;
      } else if (s[e] == '"' || s[e] == '\'') {
        q = s[e];
      }
      ++e;
    }
    if (e >= s.size()) throw std::runtime_error("unterminated XML tag");
    t = xml_tag{};
    t.begin = a;
    t.end = e + 1;
    size_t i = a + 1;
    if (s[i] == '/') {
      t.closing = true;
      ++i;
    }
    size_t n0 = i;
    while (i < e && !isspace((unsigned char)s[i]) && s[i] != '/') ++i;
    t.name = s.substr(n0, i - n0);
    for (;;) {
      while (i < e && isspace((unsigned char)s[i])) ++i;
      if (i >= e) break;
      if (s[i] == '/') {
        t.self_closing = true;
        ++i;
        continue;
      }
      size_t k0 = i;
      while (i < e && s[i] != '=' && !isspace((unsigned char)s[i])) ++i;
      std::string key = s.substr(k0, i - k0);
      while (i < e && (isspace((unsigned char)s[i]) || s[i] == '=')) ++i;
      if (i < e && (s[i] == '"' || s[i] == '\'')) {
        char qq = s[i++];
        size_t v0 = i;
        while (i < e && s[i] != qq) ++i;
        t.attr[key] = s.substr(v0, i - v0);
        ++i;
      } else {
        t.attr[key] = "";
      }
    }
    pos = t.end;
    return true;
  }
}

// Base64 decoding that continues across padded segments: VTK encodes the
// header and the (compressed) data of one array as separate segments.
class b64_reader {
 public:
  b64_reader(const std::string& s, size_t pos) : s_(s), pos_(pos) {}
  // Most bytes the rest of the text can still decode to.
  size_t max_remaining() const { return (size_t)have_ + (s_.size() - std::min(pos_, s_.size())) / 4 * 3 + 3; }
  std::vector<unsigned char> read(size_t nbytes) {
    if (nbytes > max_remaining()) throw std::runtime_error("base64 data ends early");
    std::vector<unsigned char> out;
    out.reserve(nbytes);
    while (out.size() < nbytes) {
      if (have_ == 0 && !refill()) throw std::runtime_error("base64 data ends early");
      out.push_back(buf_[idx_++]);
      --have_;
    }
    return out;
  }
 private:
  static int val(char c) {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+') return 62;
    if (c == '/') return 63;
    if (c == '=') return -2;
    return -1;
  }
  bool refill() {
    int v[4];
    int k = 0;
    while (k < 4) {
      if (pos_ >= s_.size()) return false;
      int x = val(s_[pos_++]);
      if (x == -1) {
        if (s_[pos_ - 1] == '<') return false;
        continue;  // whitespace
      }
      v[k++] = x;
    }
    idx_ = 0;
    if (v[2] == -2) {
      buf_[0] = (unsigned char)((v[0] << 2) | (v[1] >> 4));
      have_ = 1;
    } else if (v[3] == -2) {
      buf_[0] = (unsigned char)((v[0] << 2) | (v[1] >> 4));
      buf_[1] = (unsigned char)(((v[1] & 15) << 4) | (v[2] >> 2));
      have_ = 2;
    } else {
      buf_[0] = (unsigned char)((v[0] << 2) | (v[1] >> 4));
      buf_[1] = (unsigned char)(((v[1] & 15) << 4) | (v[2] >> 2));
      buf_[2] = (unsigned char)(((v[2] & 3) << 6) | v[3]);
      have_ = 3;
    }
    return true;
  }
  const std::string& s_;
  size_t pos_;
  unsigned char buf_[3] = {0, 0, 0};
  int have_ = 0, idx_ = 0;
};

// Byte source for binary array data: base64 text (inline or appended) or raw
// appended bytes.
struct byte_source {
  const std::string* s = nullptr;
  size_t pos = 0;
  bool base64 = true;
  b64_reader* b64 = nullptr;
  // Most bytes the source can still deliver.
  size_t max_remaining() const { return base64 ? b64->max_remaining() : s->size() - std::min(pos, s->size()); }
  std::vector<unsigned char> read(size_t n) {
    if (base64) return b64->read(n);
    if (pos > s->size() || n > s->size() - pos) throw std::runtime_error("raw appended data ends early");
    std::vector<unsigned char> out(s->begin() + pos, s->begin() + pos + n);
    pos += n;
    return out;
  }
};

inline uint64_t read_uint(const unsigned char* p, int size) {
  uint64_t v = 0;
  std::memcpy(&v, p, size);  // little endian (checked at load)
  return v;
}

inline size_t type_size(const std::string& t) {
  if (t == "Float32" || t == "Int32" || t == "UInt32") return 4;
  if (t == "Float64" || t == "Int64" || t == "UInt64") return 8;
  if (t == "Int16" || t == "UInt16") return 2;
  if (t == "Int8" || t == "UInt8") return 1;
  throw std::runtime_error("unsupported VTK data type " + t);
}

inline double convert(const unsigned char* p, const std::string& t) {
  if (t == "Float32") { float v; std::memcpy(&v, p, 4); return v; }
  if (t == "Float64") { double v; std::memcpy(&v, p, 8); return v; }
  if (t == "Int32") { int32_t v; std::memcpy(&v, p, 4); return v; }
  if (t == "UInt32") { uint32_t v; std::memcpy(&v, p, 4); return v; }
  if (t == "Int64") { int64_t v; std::memcpy(&v, p, 8); return (double)v; }
  if (t == "UInt64") { uint64_t v; std::memcpy(&v, p, 8); return (double)v; }
  if (t == "Int16") { int16_t v; std::memcpy(&v, p, 2); return v; }
  if (t == "UInt16") { uint16_t v; std::memcpy(&v, p, 2); return v; }
  if (t == "Int8") { int8_t v; std::memcpy(&v, p, 1); return v; }
  return (double)*p;
}

// Header + payload of one binary array (uncompressed: [nbytes] data;
// zlib: [nblocks, block size, last block size, compressed sizes...] blocks).
// A padded base64 group ends a segment; the reader then continues with the
// next one, so header and data may be encoded together or separately.
inline std::vector<unsigned char> read_binary_block(byte_source& src, int hsize, bool zlib_data) {
  if (!zlib_data) {
    auto h = src.read(hsize);
    const uint64_t nbytes = read_uint(h.data(), hsize);
    return src.read(nbytes);
  }
  auto h0 = src.read(hsize);
  const uint64_t nblocks = read_uint(h0.data(), hsize);
  // bound the block count by what the data can hold before sizing anything by it
  if (nblocks > src.max_remaining() / (uint64_t)hsize) throw std::runtime_error("zlib header ends early");
  auto h1 = src.read((2 + nblocks) * hsize);
  const uint64_t bsize = read_uint(h1.data(), hsize);
  const uint64_t last = read_uint(h1.data() + hsize, hsize);
  std::vector<uint64_t> csize(nblocks);
  for (uint64_t b = 0; b < nblocks; ++b) csize[b] = read_uint(h1.data() + (2 + b) * hsize, hsize);
  std::vector<unsigned char> out;
  for (uint64_t b = 0; b < nblocks; ++b) {
    auto c = src.read(csize[b]);
    uLongf usize = (b + 1 == nblocks && last != 0) ? last : bsize;
    const size_t at = out.size();
    out.resize(at + usize);
    if (uncompress(out.data() + at, &usize, c.data(), (uLong)c.size()) != Z_OK)
      throw std::runtime_error("zlib block failed to decompress");
    out.resize(at + usize);
  }
  return out;
}

struct data_array {
  std::string name, type, format;
  int ncomp = 1;
  std::vector<double> values;
};

struct polydata {
  uint64_t n_points = 0, n_verts = 0, n_polys = 0;
  std::vector<double> points;  // 3 per point
  std::vector<data_array> cell_data;
  std::string cell_normals_name;
  const data_array* cell_array(const std::string& name) const {
    for (auto& a : cell_data)
      if (a.name == name) return &a;
    return nullptr;
  }
};

inline polydata read_vtp(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::string s((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  // appended data (possibly raw bytes) is not XML: the tags end at it
  size_t limit = s.size(), app_data = std::string::npos;
  bool app_base64 = true;
  {
    size_t a = s.find("<AppendedData");
    if (a != std::string::npos) {
      size_t p = a;
      xml_tag t;
      next_tag(s, p, t, s.size());
      app_base64 = t.attr.count("encoding") == 0 || t.attr["encoding"] == "base64";
      size_t u = s.find('_', t.end);
      if (u == std::string::npos) throw std::runtime_error("AppendedData without '_'");
      app_data = u + 1;
      limit = a;
    }
  }
  polydata pd;
  int hsize = 4;
  bool zlib_data = false;
  size_t pos = 0;
  xml_tag t;
  std::string section;  // Points, CellData, PointData, Verts, ...
  while (next_tag(s, pos, t, limit)) {
    if (t.closing) {
      if (t.name == section) section.clear();
      continue;
    }
    if (t.name == "VTKFile") {
      if (t.attr.count("type") && t.attr["type"] != "PolyData")
        throw std::runtime_error("not a PolyData file: " + t.attr["type"]);
      if (t.attr.count("byte_order") && t.attr["byte_order"] != "LittleEndian")
        throw std::runtime_error("big-endian VTK files are not supported");
      if (t.attr.count("header_type") && t.attr["header_type"] == "UInt64") hsize = 8;
      if (t.attr.count("compressor") && !t.attr["compressor"].empty()) {
        if (t.attr["compressor"] != "vtkZLibDataCompressor")
          throw std::runtime_error("unsupported compressor " + t.attr["compressor"]);
        zlib_data = true;
      }
    } else if (t.name == "Piece") {
      pd.n_points = std::stoull(t.attr["NumberOfPoints"]);
      if (t.attr.count("NumberOfVerts")) pd.n_verts = std::stoull(t.attr["NumberOfVerts"]);
      if (t.attr.count("NumberOfPolys")) pd.n_polys = std::stoull(t.attr["NumberOfPolys"]);
    } else if (t.name == "Points" || t.name == "CellData" || t.name == "PointData" || t.name == "Verts" ||
               t.name == "Polys" || t.name == "Lines" || t.name == "Strips" || t.name == "FieldData") {
      if (!t.self_closing) section = t.name;
      if (t.name == "CellData" && t.attr.count("Normals")) pd.cell_normals_name = t.attr["Normals"];
    } else if (t.name == "DataArray") {
      data_array a;
      a.name = t.attr["Name"];
      a.type = t.attr["type"];
      a.format = t.attr.count("format") ? t.attr["format"] : "ascii";
      if (t.attr.count("NumberOfComponents")) a.ncomp = std::stoi(t.attr["NumberOfComponents"]);
      const bool wanted = section == "Points" || section == "CellData";
      if (wanted) {
        if (a.format == "ascii") {
          size_t e = s.find("</DataArray>", t.end);
          std::istringstream in(s.substr(t.end, e - t.end));
          double v;
          while (in >> v) a.values.push_back(v);
        } else {
          const size_t ts = type_size(a.type);
          std::vector<unsigned char> bytes;
          if (a.format == "binary") {
            b64_reader r(s, t.end);
            byte_source src{&s, t.end, true, &r};
            bytes = read_binary_block(src, hsize, zlib_data);
          } else if (a.format == "appended") {
            if (app_data == std::string::npos) throw std::runtime_error("appended array without AppendedData");
            const size_t off = app_data + std::stoull(t.attr["offset"]);
            b64_reader r(s, off);
            byte_source src{&s, off, app_base64, &r};
            bytes = read_binary_block(src, hsize, zlib_data);
          } else {
            throw std::runtime_error("unsupported DataArray format " + a.format);
          }
          a.values.resize(bytes.size() / ts);
          for (size_t k = 0; k < a.values.size(); ++k) a.values[k] = convert(bytes.data() + k * ts, a.type);
        }
      }
      if (section == "Points") pd.points = std::move(a.values);
      else if (section == "CellData") pd.cell_data.push_back(std::move(a));
      if (!t.self_closing) {
        size_t e = s.find("</DataArray>", t.end);
        if (e != std::string::npos && e < limit) pos = e + 12;
      }
    }
  }
  if (pd.points.size() != 3 * pd.n_points) throw std::runtime_error("point array does not match NumberOfPoints");
  return pd;
}

struct disc_cloud {
  std::vector<float> xyzr;  // 4 per disc
  std::vector<float> nrm;   // 3 per disc
  std::vector<std::string> warnings;
};

// Discs from a point cloud (vtp_point_cloud_reader.hpp:23-95).
inline disc_cloud discs_from(const polydata& pd) {
  disc_cloud dc;
  const uint64_t n = pd.n_points;
  const uint64_t ncells = pd.n_verts + pd.n_polys;
  if (n != ncells)
    dc.warnings.push_back("expects an input such that number of points equals number of cells");
  const data_array* radius = pd.cell_array("radius");
  const data_array* sqa = pd.cell_array("sqrt-of-area");
  if (!radius && !sqa) throw std::runtime_error("no cell data named \"radius\" or \"sqrt-of-area\"");
  const data_array* normals = pd.cell_normals_name.empty() ? nullptr : pd.cell_array(pd.cell_normals_name);
  if (!normals) normals = pd.cell_array("Normals");
  if (!normals) normals = pd.cell_array("normals");
  if (!normals)
    for (auto& a : pd.cell_data)
      if (a.ncomp == 3) { normals = &a; break; }
  if (!normals) throw std::runtime_error("no surface normal cell data");
  const data_array* rad = radius ? radius : sqa;
  if (rad->values.size() < n || normals->values.size() < 3 * n)
    throw std::runtime_error("cell data shorter than the number of points");
  dc.xyzr.resize(4 * n);
  dc.nrm.resize(3 * n);
  for (uint64_t i = 0; i < n; ++i) {
    double r = rad->values[i];
    if (!radius) r *= std::sqrt(3.0) / 2 * (1 + 1.0 / 32);  // radiusEpsilon, about 3 %
    dc.xyzr[4 * i + 0] = (float)pd.points[3 * i + 0];
    dc.xyzr[4 * i + 1] = (float)pd.points[3 * i + 1];
    dc.xyzr[4 * i + 2] = (float)pd.points[3 * i + 2];
    dc.xyzr[4 * i + 3] = (float)r;
    float v[3] = {(float)normals->values[3 * i], (float)normals->values[3 * i + 1], (float)normals->values[3 * i + 2]};
    const float len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (!(1 - 1e-6f <= len && len <= 1 + 1e-6f))
      for (float& x : v) x /= len;
    for (int k = 0; k < 3; ++k) dc.nrm[3 * i + k] = v[k];
  }
  return dc;
}

// ---- writing (ascii) ----
inline void put(std::ostream& o, double v) {
  char b[32];
  std::snprintf(b, sizeof b, "%.17g", v);
  o << b;
}
inline void put(std::ostream& o, float v) {
  char b[32];
  std::snprintf(b, sizeof b, "%.9g", (double)v);
  o << b;
}

struct cell_array_out {
  std::string name, type;  // "Float64", "Float32", "UInt32"
  int ncomp = 1;
  std::vector<double> values;
};

inline void write_array(std::ostream& o, const cell_array_out& a) {
  o << "        <DataArray type=\"" << a.type << "\" Name=\"" << a.name << "\"";
  if (a.ncomp != 1) o << " NumberOfComponents=\"" << a.ncomp << "\"";
  o << " format=\"ascii\">\n";
  for (size_t i = 0; i < a.values.size(); ++i) {
    o << (i % 6 == 0 ? "          " : " ");
    if (a.type == "Float32") put(o, (float)a.values[i]);
    else if (a.type == "Float64") put(o, a.values[i]);
    else o << (unsigned long long)a.values[i];
    if (i % 6 == 5 || i + 1 == a.values.size()) o << "\n";
  }
  o << "        </DataArray>\n";
}

// String field data: the characters' codes, each string ended by 0.
inline void write_field(std::ostream& o, const std::vector<std::pair<std::string, std::string>>& meta) {
  if (meta.empty()) return;
  o << "    <FieldData>\n";
  for (auto& kv : meta) {
    o << "      <Array type=\"String\" Name=\"" << kv.first << "\" NumberOfTuples=\"1\" format=\"ascii\">\n        ";
    for (unsigned char c : kv.second) o << (int)c << " ";
    o << "0\n      </Array>\n";
  }
  o << "    </FieldData>\n";
}

// Point cloud: one vertex cell per point; the normals array is the cell
// normal attribute.
inline void write_point_cloud(const std::string& path, const std::vector<float>& xyzr, const std::vector<float>& nrm,
                              const std::vector<cell_array_out>& arrays,
                              const std::vector<std::pair<std::string, std::string>>& meta) {
  const size_t n = xyzr.size() / 4;
  std::ofstream o(path);
  if (!o) throw std::runtime_error("cannot write " + path);
  o << "<?xml version=\"1.0\"?>\n<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"LittleEndian\" "
       "header_type=\"UInt64\">\n  <PolyData>\n";
  write_field(o, meta);
  o << "    <Piece NumberOfPoints=\"" << n << "\" NumberOfVerts=\"" << n
    << "\" NumberOfLines=\"0\" NumberOfStrips=\"0\" NumberOfPolys=\"0\">\n";
  o << "      <PointData>\n      </PointData>\n      <CellData Normals=\"Normals\">\n";
  cell_array_out nn{"Normals", "Float64", 3, std::vector<double>(nrm.begin(), nrm.end())};
  write_array(o, nn);
  cell_array_out rr{"radius", "Float64", 1, {}};
  rr.values.resize(n);
  for (size_t i = 0; i < n; ++i) rr.values[i] = xyzr[4 * i + 3];
  write_array(o, rr);
  for (auto& a : arrays) write_array(o, a);
  o << "      </CellData>\n      <Points>\n";
  cell_array_out pts{"Points", "Float32", 3, {}};
  pts.values.resize(3 * n);
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) pts.values[3 * i + k] = xyzr[4 * i + k];
  write_array(o, pts);
  o << "      </Points>\n      <Verts>\n        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
  for (size_t i = 0; i < n; ++i) o << (i % 12 == 0 ? "          " : " ") << i << (i % 12 == 11 || i + 1 == n ? "\n" : "");
  o << "        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
  for (size_t i = 0; i < n; ++i)
    o << (i % 12 == 0 ? "          " : " ") << i + 1 << (i % 12 == 11 || i + 1 == n ? "\n" : "");
  o << "        </DataArray>\n      </Verts>\n    </Piece>\n  </PolyData>\n</VTKFile>\n";
}

// Triangle surface with one normal per triangle.
inline void write_triangles(const std::string& path, const std::vector<double>& verts,
                            const std::vector<uint64_t>& tris, const std::vector<double>& normals) {
  const size_t nv = verts.size() / 3, nt = tris.size() / 3;
  std::ofstream o(path);
  if (!o) throw std::runtime_error("cannot write " + path);
  o << "<?xml version=\"1.0\"?>\n<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"LittleEndian\" "
       "header_type=\"UInt64\">\n  <PolyData>\n";
  o << "    <Piece NumberOfPoints=\"" << nv << "\" NumberOfVerts=\"0\" NumberOfLines=\"0\" NumberOfStrips=\"0\" "
    << "NumberOfPolys=\"" << nt << "\">\n";
  o << "      <CellData Normals=\"Normals\">\n";
  write_array(o, cell_array_out{"Normals", "Float64", 3, normals});
  o << "      </CellData>\n      <Points>\n";
  write_array(o, cell_array_out{"Points", "Float32", 3, verts});
  o << "      </Points>\n      <Polys>\n        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
  for (size_t i = 0; i < tris.size(); ++i) o << (i % 3 == 0 ? "          " : " ") << tris[i] << (i % 3 == 2 ? "\n" : "");
  o << "        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
  for (size_t i = 0; i < nt; ++i) o << "          " << 3 * (i + 1) << "\n";
  o << "        </DataArray>\n      </Polys>\n    </Piece>\n  </PolyData>\n</VTKFile>\n";
}

}  // namespace rti_vtk
