#include "headless/mesh_io.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace silk::headless {
namespace {

std::string lower_ext(const std::string& path) {
  const size_t dot = path.find_last_of('.');
  std::string e = dot == std::string::npos ? "" : path.substr(dot + 1);
  for (auto& c : e) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return e;
}

std::string read_all(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open mesh file '" + path + "'");
  std::ostringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

void add_polygon(TriMesh& m, const std::vector<int>& poly) {
  for (size_t k = 1; k + 1 < poly.size(); ++k)
    m.F.insert(m.F.end(), {poly[0], poly[k], poly[k + 1]});
}

TriMesh read_obj(const std::string& text) {
  TriMesh m;
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string tag;
    if (!(ls >> tag)) continue;
    if (tag == "v") {
      float x = 0, y = 0, z = 0;
      ls >> x >> y >> z;
      m.V.insert(m.V.end(), {x, y, z});
    } else if (tag == "f") {
      std::vector<int> poly;
      std::string tok;
      const int nv = static_cast<int>(m.V.size() / 3);
      while (ls >> tok) {
        const int idx = std::stoi(tok.substr(0, tok.find('/')));
        poly.push_back(idx < 0 ? nv + idx : idx - 1);
      }
      add_polygon(m, poly);
    }
  }
  return m;
}

TriMesh read_off(const std::string& text) {
  std::istringstream in(text);
  std::string head;
  in >> head;
  if (head.rfind("OFF", 0) != 0) throw std::runtime_error("not an OFF file");
  // Comments may follow the header; skip them.
  auto next_token = [&](std::string& t) {
    while (in >> t) {
      if (t[0] != '#') return true;
      std::string rest;
      std::getline(in, rest);
    }
    return false;
  };
  std::string t;
  next_token(t);
  const int nv = std::stoi(t);
  next_token(t);
  const int nf = std::stoi(t);
  next_token(t);  // edge count
  TriMesh m;
  for (int i = 0; i < 3 * nv; ++i) {
    next_token(t);
    m.V.push_back(std::stof(t));
  }
  for (int f = 0; f < nf; ++f) {
    next_token(t);
    const int k = std::stoi(t);
    std::vector<int> poly(k);
    for (auto& p : poly) {
      next_token(t);
      p = std::stoi(t);
    }
    std::string rest;
    std::getline(in, rest);  // optional colour
    add_polygon(m, poly);
  }
  return m;
}

// PLY property types.
int ply_size(const std::string& t) {
  if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
  if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
  if (t == "int" || t == "uint" || t == "float" || t == "int32" || t == "uint32" ||
      t == "float32")
    return 4;
  if (t == "double" || t == "float64") return 8;
  throw std::runtime_error("PLY: unknown property type '" + t + "'");
}

double ply_read(const char* p, const std::string& t) {
  if (t == "char" || t == "int8") return static_cast<double>(*reinterpret_cast<const int8_t*>(p));
  if (t == "uchar" || t == "uint8") return static_cast<double>(*reinterpret_cast<const uint8_t*>(p));
  int16_t i16;
  uint16_t u16;
  int32_t i32;
  uint32_t u32;
  float f;
  double d;
  if (t == "short" || t == "int16") return std::memcpy(&i16, p, 2), static_cast<double>(i16);
  if (t == "ushort" || t == "uint16") return std::memcpy(&u16, p, 2), static_cast<double>(u16);
  if (t == "int" || t == "int32") return std::memcpy(&i32, p, 4), static_cast<double>(i32);
  if (t == "uint" || t == "uint32") return std::memcpy(&u32, p, 4), static_cast<double>(u32);
  if (t == "float" || t == "float32") return std::memcpy(&f, p, 4), static_cast<double>(f);
  return std::memcpy(&d, p, 8), d;
}

struct PlyProp {
  std::string name, type, count_type;  // count_type set for lists
};
struct PlyElem {
  std::string name;
  size_t count = 0;
  std::vector<PlyProp> props;
};

TriMesh read_ply(const std::string& text) {
  const size_t end = text.find("end_header");
  if (text.rfind("ply", 0) != 0 || end == std::string::npos)
    throw std::runtime_error("not a PLY file");
  std::istringstream hs(text.substr(0, end));
  std::string line, format;
  std::vector<PlyElem> elems;
  while (std::getline(hs, line)) {
    std::istringstream ls(line);
    std::string w;
    ls >> w;
    if (w == "format") {
      ls >> format;
    } else if (w == "element") {
      PlyElem e;
      ls >> e.name >> e.count;
      elems.push_back(e);
    } else if (w == "property" && !elems.empty()) {
      PlyProp p;
      ls >> p.type;
      if (p.type == "list") {
        ls >> p.count_type >> p.type >> p.name;
      } else {
        ls >> p.name;
      }
      elems.back().props.push_back(p);
    }
  }
  size_t body = text.find('\n', end);
  if (body == std::string::npos) throw std::runtime_error("PLY: no body");
  ++body;
  TriMesh m;
  if (format == "ascii") {
    std::istringstream bs(text.substr(body));
    for (const auto& e : elems)
      for (size_t r = 0; r < e.count; ++r) {
        std::array<double, 3> xyz{};
        for (const auto& p : e.props) {
          if (!p.count_type.empty()) {
            int k = 0;
            bs >> k;
            std::vector<int> poly(k);
            for (auto& q : poly) bs >> q;
            if (e.name == "face") add_polygon(m, poly);
          } else {
            double v = 0;
            bs >> v;
            if (e.name == "vertex") {
              if (p.name == "x") xyz[0] = v;
              if (p.name == "y") xyz[1] = v;
              if (p.name == "z") xyz[2] = v;
            }
          }
        }
        if (e.name == "vertex")
          m.V.insert(m.V.end(), {static_cast<float>(xyz[0]), static_cast<float>(xyz[1]),
                                 static_cast<float>(xyz[2])});
      }
    return m;
  }
  if (format != "binary_little_endian")
    throw std::runtime_error("PLY: unsupported format '" + format + "'");
  const char* p = text.data() + body;
  const char* stop = text.data() + text.size();
  auto need = [&](size_t n) {
    if (p + n > stop) throw std::runtime_error("PLY: truncated body");
  };
  for (const auto& e : elems)
    for (size_t r = 0; r < e.count; ++r) {
      std::array<double, 3> xyz{};
      for (const auto& pr : e.props) {
        if (!pr.count_type.empty()) {
          need(ply_size(pr.count_type));
          const int k = static_cast<int>(ply_read(p, pr.count_type));
          p += ply_size(pr.count_type);
          std::vector<int> poly(k);
          for (auto& q : poly) {
            need(ply_size(pr.type));
            q = static_cast<int>(ply_read(p, pr.type));
            p += ply_size(pr.type);
          }
          if (e.name == "face") add_polygon(m, poly);
        } else {
          need(ply_size(pr.type));
          const double v = ply_read(p, pr.type);
          p += ply_size(pr.type);
          if (e.name == "vertex") {
            if (pr.name == "x") xyz[0] = v;
            if (pr.name == "y") xyz[1] = v;
            if (pr.name == "z") xyz[2] = v;
          }
        }
      }
      if (e.name == "vertex")
        m.V.insert(m.V.end(), {static_cast<float>(xyz[0]), static_cast<float>(xyz[1]),
                               static_cast<float>(xyz[2])});
    }
  return m;
}

TriMesh read_stl(const std::string& text) {
  std::vector<std::array<float, 9>> tris;
  const bool ascii = text.rfind("solid", 0) == 0 && text.find("facet") != std::string::npos;
  if (ascii) {
    std::istringstream in(text);
    std::string w;
    std::array<float, 9> t{};
    int k = 0;
    while (in >> w)
      if (w == "vertex") {
        in >> t[3 * k] >> t[3 * k + 1] >> t[3 * k + 2];
        if (++k == 3) {
          tris.push_back(t);
          k = 0;
        }
      }
  } else {
    if (text.size() < 84) throw std::runtime_error("STL: truncated header");
    uint32_t n = 0;
    std::memcpy(&n, text.data() + 80, 4);
    if (text.size() < 84 + 50 * static_cast<size_t>(n)) throw std::runtime_error("STL: truncated");
    for (uint32_t i = 0; i < n; ++i) {
      std::array<float, 9> t;
      std::memcpy(t.data(), text.data() + 84 + 50 * static_cast<size_t>(i) + 12, 36);
      tris.push_back(t);
    }
  }
  TriMesh m;
  std::map<std::array<float, 3>, int> ids;
  for (const auto& t : tris)
    for (int k = 0; k < 3; ++k) {
      const std::array<float, 3> p{t[3 * k], t[3 * k + 1], t[3 * k + 2]};
      auto [it, fresh] = ids.emplace(p, static_cast<int>(m.V.size() / 3));
      if (fresh) m.V.insert(m.V.end(), p.begin(), p.end());
      m.F.push_back(it->second);
    }
  return m;
}

}  // namespace

TriMesh read_mesh(const std::string& path) {
  const std::string ext = lower_ext(path);
  const std::string text = read_all(path);
  TriMesh m;
  if (ext == "obj") m = read_obj(text);
  else if (ext == "off") m = read_off(text);
  else if (ext == "ply") m = read_ply(text);
  else if (ext == "stl") m = read_stl(text);
  else throw std::runtime_error("unsupported mesh format '." + ext + "'");
  const int nv = static_cast<int>(m.V.size() / 3);
  for (int i : m.F)
    if (i < 0 || i >= nv) throw std::runtime_error("mesh '" + path + "': face index out of range");
  return m;
}

}  // namespace silk::headless
