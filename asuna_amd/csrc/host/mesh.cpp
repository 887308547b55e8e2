// Wavefront OBJ loading with the semantics of loadMesh (src/core/mesh.cpp:112-145):
// every face corner becomes its own GpuVertex (de-indexed), the index buffer
// is the running corner count, uv.y is flipped (1 - v), missing texcoords or
// normals read as zero.  Polygons are fan-triangulated (v0, vi, vi+1).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "scene.h"

namespace ash {

namespace {

struct Corner {
  long v, t, n;  // 0-based, -1 = absent
};

inline const char* skipWs(const char* p) {
  while (*p == ' ' || *p == '\t') ++p;
  return p;
}

inline bool parseFloat(const char*& p, float& out) {
  p = skipWs(p);
  char* e;
  out = std::strtof(p, &e);
  if (e == p) return false;
  p = e;
  return true;
}

// OBJ index: 1-based, negative = relative to the current count.
inline long fixIndex(long i, size_t count) {
  if (i > 0) return i - 1;
  if (i < 0) return long(count) + i;
  return -1;
}

bool parseCorner(const char*& p, size_t nv, size_t nt, size_t nn, Corner& c) {
  p = skipWs(p);
  if (*p == '\0' || *p == '\n' || *p == '\r' || *p == '#') return false;
  char* e;
  long v = std::strtol(p, &e, 10);
  if (e == p) return false;
  p = e;
  c.v = fixIndex(v, nv);
  c.t = c.n = -1;
  if (*p == '/') {
    ++p;
    if (*p != '/') {
      long t = std::strtol(p, &e, 10);
      if (e != p) c.t = fixIndex(t, nt);
      p = e;
    }
    if (*p == '/') {
      ++p;
      long n = std::strtol(p, &e, 10);
      if (e != p) c.n = fixIndex(n, nn);
      p = e;
    }
  }
  while (*p && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') ++p;
  return true;
}

}  // namespace

void loadObj(const std::string& path, std::vector<GpuVertex>& out, std::vector<uint32_t>& idx) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("Scene: load mesh from [" + path + "] --- cannot open file");
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::string buf(size_t(sz) + 1, '\0');
  if (sz > 0 && std::fread(&buf[0], 1, size_t(sz), f) != size_t(sz)) {
    std::fclose(f);
    throw std::runtime_error("Scene: load mesh from [" + path + "] --- read error");
  }
  std::fclose(f);
  std::vector<float> pos, tex, nrm;
  std::vector<Corner> face;
  const char* p = buf.c_str();
  size_t line = 0;
  while (*p) {
    ++line;
    const char* q = skipWs(p);
    if (q[0] == 'v' && (q[1] == ' ' || q[1] == '\t')) {
      q += 2;
      float a, b, c;
      if (!parseFloat(q, a) || !parseFloat(q, b) || !parseFloat(q, c))
        throw std::runtime_error("Scene: load mesh from [" + path + "] --- bad vertex at line " + std::to_string(line));
      pos.push_back(a), pos.push_back(b), pos.push_back(c);
    } else if (q[0] == 'v' && q[1] == 't' && (q[2] == ' ' || q[2] == '\t')) {
      q += 3;
      float a = 0, b = 0;
      if (!parseFloat(q, a)) throw std::runtime_error("Scene: bad texcoord in [" + path + "]");
      parseFloat(q, b);
      tex.push_back(a), tex.push_back(b);
    } else if (q[0] == 'v' && q[1] == 'n' && (q[2] == ' ' || q[2] == '\t')) {
      q += 3;
      float a, b, c;
      if (!parseFloat(q, a) || !parseFloat(q, b) || !parseFloat(q, c))
        throw std::runtime_error("Scene: bad normal in [" + path + "]");
      nrm.push_back(a), nrm.push_back(b), nrm.push_back(c);
    } else if (q[0] == 'f' && (q[1] == ' ' || q[1] == '\t')) {
      q += 2;
      face.clear();
      Corner c;
      while (parseCorner(q, pos.size() / 3, tex.size() / 2, nrm.size() / 3, c)) face.push_back(c);
      if (face.size() < 3) throw std::runtime_error("Scene: degenerate face in [" + path + "]");
      for (size_t k = 1; k + 1 < face.size(); ++k) {
        const Corner* tri[3] = {&face[0], &face[k], &face[k + 1]};
        for (const Corner* cc : tri) {
          GpuVertex vx;
          std::memset(&vx, 0, sizeof vx);
          if (cc->v < 0 || size_t(cc->v) * 3 + 2 >= pos.size())
            throw std::runtime_error("Scene: vertex index out of range in [" + path + "]");
          std::memcpy(vx.pos, &pos[size_t(cc->v) * 3], 12);
          if (!tex.empty() && cc->t >= 0 && size_t(cc->t) * 2 + 1 < tex.size()) {
            vx.uv[0] = tex[size_t(cc->t) * 2];
            vx.uv[1] = 1.0f - tex[size_t(cc->t) * 2 + 1];
          }
          if (!nrm.empty() && cc->n >= 0 && size_t(cc->n) * 3 + 2 < nrm.size())
            std::memcpy(vx.normal, &nrm[size_t(cc->n) * 3], 12);
          out.push_back(vx);
          idx.push_back(uint32_t(idx.size()));
        }
      }
    }
    while (*p && *p != '\n') ++p;
    if (*p == '\n') ++p;
  }
}

}  // namespace ash
