// Procedural benchmark scenes, written as real scene files (JSON + OBJ +
// EXR) in the reference's format so that the loader is part of every run.
// Seeds: master seed (default 0xA5A52026); per-object streams come from
// SplitMix64 over (config, object index).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "scene.h"

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {  // SplitMix64
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uni() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
  float uf(float a, float b) { return float(a + (b - a) * uni()); }
};
Rng objectRng(uint64_t master, int config, int object) {
  Rng r(master ^ (uint64_t(config + 1) << 40) ^ (uint64_t(object) * 0x2545F4914F6CDD1Dull));
  r.next();
  return r;
}

struct V3 {
  double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 nrm(V3 a) {
  double l = std::sqrt(dot(a, a));
  return l > 0 ? a * (1.0 / l) : a;
}

struct Mesh {
  std::vector<V3> v, n;
  std::vector<int> f;  // 3 per triangle, 0-based
};

bool writeObj(const std::string& path, const Mesh& m) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) return false;
  std::vector<char> buf(1 << 20);
  std::setvbuf(fp, buf.data(), _IOFBF, buf.size());
  for (auto& p : m.v) std::fprintf(fp, "v %.9g %.9g %.9g\n", p.x, p.y, p.z);
  for (auto& p : m.n) std::fprintf(fp, "vn %.9g %.9g %.9g\n", p.x, p.y, p.z);
  bool hasN = !m.n.empty();
  for (size_t i = 0; i < m.f.size(); i += 3) {
    int a = m.f[i] + 1, b = m.f[i + 1] + 1, c = m.f[i + 2] + 1;
    if (hasN) std::fprintf(fp, "f %d//%d %d//%d %d//%d\n", a, a, b, b, c, c);
    else std::fprintf(fp, "f %d %d %d\n", a, b, c);
  }
  return std::fclose(fp) == 0;
}

// Axis-aligned box [-0.5,0.5]^3 as 12 triangles with flat outward normals
// (24 vertices).
Mesh unitBox() {
  Mesh m;
  const int faces[6][4][3] = {
      {{1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}}, {{0, 0, 1}, {0, 1, 1}, {0, 1, 0}, {0, 0, 0}},
      {{0, 1, 0}, {0, 1, 1}, {1, 1, 1}, {1, 1, 0}}, {{0, 0, 1}, {0, 0, 0}, {1, 0, 0}, {1, 0, 1}},
      {{0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}}, {{1, 0, 0}, {0, 0, 0}, {0, 1, 0}, {1, 1, 0}}};
  const V3 nn[6] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  for (int f = 0; f < 6; ++f) {
    int base = int(m.v.size());
    for (int k = 0; k < 4; ++k) {
      m.v.push_back({faces[f][k][0] - 0.5, faces[f][k][1] - 0.5, faces[f][k][2] - 0.5});
      m.n.push_back(nn[f]);
    }
    // orient (0,1,2),(0,2,3) so that the geometric normal matches nn[f]
    V3 g = cross(m.v[base + 1] - m.v[base], m.v[base + 2] - m.v[base]);
    if (dot(g, nn[f]) >= 0) {
      m.f.insert(m.f.end(), {base, base + 1, base + 2, base, base + 2, base + 3});
    } else {
      m.f.insert(m.f.end(), {base, base + 2, base + 1, base, base + 3, base + 2});
    }
  }
  return m;
}

Mesh quad(V3 p, V3 u, V3 v, V3 n) {
  Mesh m;
  m.v = {p, p + u, p + u + v, p + v};
  m.n = {n, n, n, n};
  m.f = {0, 1, 2, 0, 2, 3};
  return m;
}

// ---- value-noise fBm (hash lattice, quintic fade), seeded per object
struct Noise {
  uint32_t seed;
  static uint32_t hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
  }
  double lattice(int x, int y, int z) const {
    uint32_t h = hash(uint32_t(x) * 73856093u ^ hash(uint32_t(y) * 19349663u ^ hash(uint32_t(z) * 83492791u ^ seed)));
    return double(h) / 4294967295.0 * 2.0 - 1.0;
  }
  static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
  double value(V3 p) const {
    int x0 = int(std::floor(p.x)), y0 = int(std::floor(p.y)), z0 = int(std::floor(p.z));
    double fx = fade(p.x - x0), fy = fade(p.y - y0), fz = fade(p.z - z0);
    double c[2][2][2];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        for (int k = 0; k < 2; ++k) c[i][j][k] = lattice(x0 + i, y0 + j, z0 + k);
    auto lerp = [](double a, double b, double t) { return a + (b - a) * t; };
    double x00 = lerp(c[0][0][0], c[1][0][0], fx), x10 = lerp(c[0][1][0], c[1][1][0], fx);
    double x01 = lerp(c[0][0][1], c[1][0][1], fx), x11 = lerp(c[0][1][1], c[1][1][1], fx);
    return lerp(lerp(x00, x10, fy), lerp(x01, x11, fy), fz);
  }
  double fbm(V3 p, int octaves) const {
    double sum = 0, amp = 0.5, norm = 0, f = 1;
    for (int o = 0; o < octaves; ++o) {
      sum += amp * value(p * f);
      norm += amp;
      amp *= 0.5;
      f *= 2.0;
    }
    return sum / norm;
  }
};

// Icosphere of `level` subdivisions (20 * 4^level faces), unit radius.
Mesh icosphere(int level) {
  Mesh m;
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  V3 iv[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
               {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  for (auto& p : iv) m.v.push_back(nrm(p));
  m.f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
         3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
  for (int l = 0; l < level; ++l) {
    std::vector<int> nf;
    nf.reserve(m.f.size() * 4);
    std::vector<std::pair<uint64_t, int>> cache;
    std::vector<std::vector<std::pair<int, int>>> edge(m.v.size());
    auto mid = [&](int a, int b) {
      int lo = std::min(a, b), hi = std::max(a, b);
      for (auto& e : edge[size_t(lo)])
        if (e.first == hi) return e.second;
      int id = int(m.v.size());
      m.v.push_back(nrm((m.v[size_t(a)] + m.v[size_t(b)]) * 0.5));
      edge[size_t(lo)].push_back({hi, id});
      edge.emplace_back();
      return id;
    };
    for (size_t i = 0; i < m.f.size(); i += 3) {
      int a = m.f[i], b = m.f[i + 1], c = m.f[i + 2];
      int ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
      nf.insert(nf.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
    }
    m.f.swap(nf);
  }
  return m;
}

void smoothNormals(Mesh& m) {
  m.n.assign(m.v.size(), V3{0, 0, 0});
  for (size_t i = 0; i < m.f.size(); i += 3) {
    int a = m.f[i], b = m.f[i + 1], c = m.f[i + 2];
    V3 g = cross(m.v[size_t(b)] - m.v[size_t(a)], m.v[size_t(c)] - m.v[size_t(a)]);  // area weighted
    for (int k : {a, b, c}) m.n[size_t(k)] = m.n[size_t(k)] + g;
  }
  for (auto& n : m.n) n = nrm(n);
}

std::string fmt3(double a, double b, double c) {
  char buf[128];
  std::snprintf(buf, sizeof buf, "[%.9g, %.9g, %.9g]", a, b, c);
  return buf;
}

bool writeText(const std::string& path, const std::string& s) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) return false;
  bool ok = std::fwrite(s.data(), 1, s.size(), fp) == s.size();
  return std::fclose(fp) == 0 && ok;
}

// Closed box (configs 0 and 2).
int genBox(uint64_t master, int config, const std::string& dir, std::string& scenePath) {
  std::string js = "{\n";
  bool multiShot = config == 2;
  int spp = multiShot ? 32 : 64, depth = multiShot ? 6 : 8;
  char buf[512];
  std::snprintf(buf, sizeof buf,
                "  \"state\": {\"path_tracing\": {\"spp\": %d, \"max_path_depth\": %d, "
                "\"multi_channel\": [\"diffuse\", \"normal\", \"position\"]}, "
                "\"post_processing\": {\"tone_mapping\": \"filmic\"}, \"output_hdr\": true},\n",
                spp, depth);
  js += buf;
  js += "  \"camera\": {\"type\": \"perspective\", \"film\": {\"resolution\": [512, 512]}, \"fov\": 40.0},\n";
  // walls: floor, ceiling, back, left, right (open towards +z)
  struct W {
    const char* name;
    V3 p, u, v, n;
  } walls[5] = {{"floor", {-1, -1, 1}, {2, 0, 0}, {0, 0, -2}, {0, 1, 0}},
                {"ceiling", {-1, 1, -1}, {2, 0, 0}, {0, 0, 2}, {0, -1, 0}},
                {"back", {-1, -1, -1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 1}},
                {"left", {-1, -1, 1}, {0, 0, -2}, {0, 2, 0}, {1, 0, 0}},
                {"right", {1, -1, -1}, {0, 0, 2}, {0, 2, 0}, {-1, 0, 0}}};
  std::string mats, meshes, insts;
  int obj = 0;
  for (auto& w : walls) {
    Mesh m = quad(w.p, w.u, w.v, w.n);
    if (!writeObj(dir + "/" + w.name + ".obj", m)) return ASH_E_IO;
    Rng r = objectRng(master, 0, obj++);
    float a = r.uf(0.1f, 0.8f), b = r.uf(0.1f, 0.8f), c = r.uf(0.1f, 0.8f);
    mats += std::string(mats.empty() ? "" : ",\n") + "    {\"name\": \"m_" + w.name +
            "\", \"type\": \"brdf_lambertian\", \"diffuse_reflectance\": " + fmt3(a, b, c) + "}";
    meshes += std::string(meshes.empty() ? "" : ",\n") + "    {\"name\": \"" + w.name + "\", \"path\": \"" + w.name + ".obj\"}";
    insts += std::string(insts.empty() ? "" : ",\n") + "    {\"mesh\": \"" + w.name + "\", \"material\": \"m_" + w.name + "\"}";
  }
  Mesh box = unitBox();
  if (!writeObj(dir + "/box.obj", box)) return ASH_E_IO;
  meshes += ",\n    {\"name\": \"box\", \"path\": \"box.obj\"}";
  const double sizes[2][3] = {{0.6, 1.2, 0.6}, {0.6, 0.6, 0.6}};
  const double pos[2][3] = {{-0.35, -0.4, -0.35}, {0.4, -0.7, 0.3}};
  for (int b = 0; b < 2; ++b) {
    Rng r = objectRng(master, 0, obj++);
    float a = r.uf(0.1f, 0.8f), g = r.uf(0.1f, 0.8f), c = r.uf(0.1f, 0.8f);
    float yaw = r.uf(0.f, 45.f);
    std::string name = "box" + std::to_string(b);
    mats += ",\n    {\"name\": \"m_" + name + "\", \"type\": \"brdf_lambertian\", \"diffuse_reflectance\": " + fmt3(a, g, c) + "}";
    insts += ",\n    {\"mesh\": \"box\", \"material\": \"m_" + name + "\", \"toworld\": [" +
             "{\"type\": \"scale\", \"value\": " + fmt3(sizes[b][0], sizes[b][1], sizes[b][2]) + "}, " +
             "{\"type\": \"roty\", \"value\": " + std::to_string(yaw) + "}, " +
             "{\"type\": \"translate\", \"value\": " + fmt3(pos[b][0], pos[b][1], pos[b][2]) + "}]}";
  }
  js += "  \"materials\": [\n" + mats + "\n  ],\n";
  js += "  \"lights\": [\n    {\"type\": \"rect\", \"radiance\": [17, 12, 4], \"position\": [-0.125, 0.999, -0.125], "
        "\"v1\": [0.125, 0.999, -0.125], \"v2\": [-0.125, 0.999, 0.125]}\n  ],\n";
  js += "  \"meshes\": [\n" + meshes + "\n  ],\n";
  js += "  \"instances\": [\n" + insts + "\n  ],\n";
  js += "  \"shots\": [\n";
  if (!multiShot) {
    js += "    {\"type\": \"lookat\", \"eye\": [0, 0, 3.4], \"lookat\": [0, 0, 0], \"up\": [0, 1, 0]}\n";
  } else {
    for (int s = 0; s < 64; ++s) {
      Rng r = objectRng(master, 2, 1000 + s);
      V3 d;
      do {
        double z = 2.0 * r.uni() - 1.0, ph = 2.0 * M_PI * r.uni(), rr = std::sqrt(std::max(0.0, 1 - z * z));
        d = {rr * std::cos(ph), z, rr * std::sin(ph)};
      } while (std::fabs(d.y) > 0.99);
      double rad = 5.0 + 4.0 * r.uni();
      double fov = 30.0 + 40.0 * r.uni();
      char sb[512];
      std::snprintf(sb, sizeof sb,
                    "    {\"type\": \"lookat\", \"eye\": %s, \"lookat\": [0, 0, 0], \"up\": [0, 1, 0], \"fov\": %.9g}%s\n",
                    fmt3(d.x * rad, d.y * rad, d.z * rad).c_str(), fov, s + 1 < 64 ? "," : "");
      js += sb;
    }
  }
  js += "  ]\n}\n";
  scenePath = dir + (multiShot ? "/multishot.json" : "/box.json");
  return writeText(scenePath, js) ? ASH_OK : ASH_E_IO;
}

// Sky: gradient + sun disk (angular diameter 0.5 deg, radiance 1e4).
void skyColor(V3 d, V3 sun, float out[3]) {
  double el = d.y;
  double r, g, b;
  if (el >= 0) {
    double t = std::pow(el, 0.5);
    r = 0.9 * (1 - t) + 0.25 * t;
    g = 0.85 * (1 - t) + 0.45 * t;
    b = 0.8 * (1 - t) + 0.95 * t;
  } else {
    r = g = b = 0.15;
  }
  double cosang = dot(d, sun);
  if (cosang > std::cos(0.25 * M_PI / 180.0)) r = g = b = 1e4;
  out[0] = float(r), out[1] = float(g), out[2] = float(b);
}

// grid x grid displaced icospheres of `level` subdivisions on a ground quad,
// envW x envW/2 sky, 4 triangle lights, film W x H.
int genSpheres(uint64_t master, int grid, int level, int envW, uint32_t W, uint32_t H, const std::string& dir,
               std::string& scenePath) {
  int envH = envW / 2;
  Mesh base = icosphere(level);
  const double radius = 0.4, spacing = 1.0;
  std::string mats, meshes, insts;
  for (int gz = 0; gz < grid; ++gz)
    for (int gx = 0; gx < grid; ++gx) {
      int id = gz * grid + gx;
      Rng r = objectRng(master, 1, id);
      Noise nz{uint32_t(r.next())};
      Mesh m = base;
      V3 off{r.uni() * 100.0, r.uni() * 100.0, r.uni() * 100.0};
      for (auto& p : m.v) {
        double disp = 0.1 * nz.fbm(p * 2.0 + off, 6);  // amplitude 0.1 r in object units (r = 1)
        p = p * (1.0 + disp);
      }
      smoothNormals(m);
      std::string name = "sphere" + std::to_string(id);
      if (!writeObj(dir + "/" + name + ".obj", m)) return ASH_E_IO;
      std::string mat;
      if (r.uni() < 0.1) {
        mat = "{\"name\": \"m_" + name + "\", \"type\": \"bsdf_dielectric\", \"ior\": 1.5}";
      } else {
        float rough = r.uf(0.05f, 1.0f);
        int metal = r.uni() < 0.5 ? 1 : 0;
        float a = r.uf(0.1f, 0.9f), b = r.uf(0.1f, 0.9f), c = r.uf(0.1f, 0.9f);
        char mb[512];
        std::snprintf(mb, sizeof mb,
                      "{\"name\": \"m_%s\", \"type\": \"brdf_pbr_metalness_roughness\", \"diffuse_reflectance\": %s, "
                      "\"roughness\": %.9g, \"metalness\": %d}",
                      name.c_str(), fmt3(a, b, c).c_str(), rough, metal);
        mat = mb;
      }
      mats += std::string(mats.empty() ? "" : ",\n") + "    " + mat;
      meshes += std::string(meshes.empty() ? "" : ",\n") + "    {\"name\": \"" + name + "\", \"path\": \"" + name + ".obj\"}";
      double cx = (gx - (grid - 1) * 0.5) * spacing, cz = (gz - (grid - 1) * 0.5) * spacing;
      insts += std::string(insts.empty() ? "" : ",\n") + "    {\"mesh\": \"" + name + "\", \"material\": \"m_" + name +
               "\", \"toworld\": [{\"type\": \"scale\", \"value\": " + fmt3(radius, radius, radius) +
               "}, {\"type\": \"translate\", \"value\": " + fmt3(cx, radius * 1.1, cz) + "}]}";
    }
  Mesh ground = quad({-20, 0, 20}, {40, 0, 0}, {0, 0, -40}, {0, 1, 0});
  if (!writeObj(dir + "/ground.obj", ground)) return ASH_E_IO;
  mats += ",\n    {\"name\": \"m_ground\", \"type\": \"brdf_lambertian\", \"diffuse_reflectance\": [0.5, 0.5, 0.5]}";
  meshes += ",\n    {\"name\": \"ground\", \"path\": \"ground.obj\"}";
  insts += ",\n    {\"mesh\": \"ground\", \"material\": \"m_ground\"}";
  // four one-triangle mesh lights facing down
  std::string lights;
  for (int l = 0; l < 4; ++l) {
    Rng r = objectRng(master, 1, 5000 + l);
    double cx = (l % 2 ? 1.5 : -1.5) + r.uf(-0.3f, 0.3f), cz = (l / 2 ? 1.5 : -1.5) + r.uf(-0.3f, 0.3f);
    Mesh t;
    t.v = {{cx - 0.3, 3.5, cz - 0.3}, {cx + 0.3, 3.5, cz - 0.3}, {cx - 0.3, 3.5, cz + 0.3}};
    t.f = {0, 1, 2};  // cross(v1-v0, v2-v0) points -y: emits downwards
    std::string name = "light" + std::to_string(l) + ".obj";
    if (!writeObj(dir + "/" + name, t)) return ASH_E_IO;
    lights += std::string(lights.empty() ? "" : ",\n") + "    {\"type\": \"mesh\", \"radiance\": [20, 18, 15], \"path\": \"" + name + "\"}";
  }
  // env map
  std::vector<float> env(size_t(envW) * envH * 4);
  V3 sun = nrm(V3{std::cos(35 * M_PI / 180) * std::cos(40 * M_PI / 180), std::sin(35 * M_PI / 180),
                  std::cos(35 * M_PI / 180) * std::sin(40 * M_PI / 180)});
  for (int j = 0; j < envH; ++j)
    for (int i = 0; i < envW; ++i) {
      double theta = (j + 0.5) / envH * M_PI, phi = (i + 0.5) / envW * 2 * M_PI - M_PI;
      V3 d{std::sin(theta) * std::cos(phi), std::cos(theta), std::sin(theta) * std::sin(phi)};
      float* o = &env[(size_t(j) * envW + i) * 4];
      skyColor(d, sun, o);
      o[3] = 1.f;
    }
  // The sun disk is narrower than a texel at 2048 columns; widen it to the
  // texel(s) its centre falls in so it is present at every resolution.
  {
    double theta = std::acos(sun.y), phi = std::atan2(sun.z, sun.x);
    int j = std::min(envH - 1, int(theta / M_PI * envH)), i = std::min(envW - 1, int((phi + M_PI) / (2 * M_PI) * envW));
    float* o = &env[(size_t(j) * envW + i) * 4];
    o[0] = o[1] = o[2] = 1e4f;
  }
  std::string err;
  if (!ash::writeImage(dir + "/sky.exr", uint32_t(envW), uint32_t(envH), env.data(), err)) return ASH_E_IO;
  std::string js = "{\n";
  js += "  \"state\": {\"path_tracing\": {\"spp\": 16, \"max_path_depth\": 6, "
        "\"multi_channel\": [\"diffuse\", \"normal\", \"depth\"]}, "
        "\"post_processing\": {\"tone_mapping\": \"filmic\"}, \"output_hdr\": true},\n";
  char cb[256];
  std::snprintf(cb, sizeof cb, "  \"camera\": {\"type\": \"perspective\", \"film\": {\"resolution\": [%u, %u]}, \"fov\": 60.0},\n", W, H);
  js += cb;
  js += "  \"materials\": [\n" + mats + "\n  ],\n";
  js += "  \"lights\": [\n" + lights + "\n  ],\n";
  js += "  \"envmap\": {\"path\": \"sky.exr\"},\n";
  js += "  \"meshes\": [\n" + meshes + "\n  ],\n";
  js += "  \"instances\": [\n" + insts + "\n  ],\n";
  double camDist = grid * 0.85 + 1.0;
  js += "  \"shots\": [\n    {\"type\": \"lookat\", \"eye\": " + fmt3(0.0, camDist * 0.55, camDist) +
        ", \"lookat\": [0, 0.3, 0], \"up\": [0, 1, 0]}\n  ]\n}\n";
  scenePath = dir + "/spheres.json";
  return writeText(scenePath, js) ? ASH_OK : ASH_E_IO;
}

// Random training scene (the denoiser's data set): 1-20 fBm-displaced
// icospheres of 320 to 5120 triangles with random size, position and
// material from the plugin set (lambertian, pbr metalness-roughness, rough
// plastic, dielectric) on a lambertian ground quad, 1-4 rect area lights
// facing down with random size and radiance, and an env map with a random
// sky gradient and sun (or a constant colour).  Every draw comes from the
// scene's own seed, so the scene set is reproducible.
int genTraining(uint64_t master, const std::string& dir, std::string& scenePath) {
  Rng r = objectRng(master, 3, 0);
  const int nObj = 1 + int(r.uni() * 20.0);
  std::string mats, meshes, insts;
  for (int id = 0; id < nObj; ++id) {
    Rng o = objectRng(master, 3, 1 + id);
    const int level = 2 + int(o.uni() * 3.0);  // 2..4
    Mesh m = icosphere(level);
    Noise nz{uint32_t(o.next())};
    V3 off{o.uni() * 100.0, o.uni() * 100.0, o.uni() * 100.0};
    const double amp = o.uf(0.02f, 0.25f);
    for (auto& p : m.v) p = p * (1.0 + amp * nz.fbm(p * 2.0 + off, 6));
    smoothNormals(m);
    std::string name = "obj" + std::to_string(id);
    if (!writeObj(dir + "/" + name + ".obj", m)) return ASH_E_IO;
    const double kind = o.uni();
    const float a = o.uf(0.05f, 0.95f), b = o.uf(0.05f, 0.95f), c = o.uf(0.05f, 0.95f);
    char mb[512];
    if (kind < 0.3) {
      std::snprintf(mb, sizeof mb, "{\"name\": \"m_%s\", \"type\": \"brdf_lambertian\", \"diffuse_reflectance\": %s}",
                    name.c_str(), fmt3(a, b, c).c_str());
    } else if (kind < 0.65) {
      std::snprintf(mb, sizeof mb,
                    "{\"name\": \"m_%s\", \"type\": \"brdf_pbr_metalness_roughness\", \"diffuse_reflectance\": %s, "
                    "\"roughness\": %.9g, \"metalness\": %.9g}",
                    name.c_str(), fmt3(a, b, c).c_str(), o.uf(0.05f, 1.0f), o.uni() < 0.5 ? 1.0 : 0.0);
    } else if (kind < 0.9) {
      std::snprintf(mb, sizeof mb,
                    "{\"name\": \"m_%s\", \"type\": \"brdf_rough_plastic\", \"diffuse_reflectance\": %s, "
                    "\"alpha\": [%.9g, %.9g], \"ior\": %.9g}",
                    name.c_str(), fmt3(a, b, c).c_str(), o.uf(0.02f, 0.6f), o.uf(0.02f, 0.6f), o.uf(1.3f, 1.8f));
    } else {
      std::snprintf(mb, sizeof mb, "{\"name\": \"m_%s\", \"type\": \"bsdf_dielectric\", \"ior\": %.9g}", name.c_str(),
                    o.uf(1.3f, 1.8f));
    }
    mats += std::string(mats.empty() ? "" : ",\n") + "    " + mb;
    meshes += std::string(meshes.empty() ? "" : ",\n") + "    {\"name\": \"" + name + "\", \"path\": \"" + name + ".obj\"}";
    const double rad = o.uf(0.15f, 0.7f);
    const double cx = o.uf(-2.0f, 2.0f), cz = o.uf(-2.0f, 2.0f), cy = rad * (1.0 + amp) + o.uf(0.0f, 0.8f);
    insts += std::string(insts.empty() ? "" : ",\n") + "    {\"mesh\": \"" + name + "\", \"material\": \"m_" + name +
             "\", \"toworld\": [{\"type\": \"scale\", \"value\": " + fmt3(rad, rad, rad) +
             "}, {\"type\": \"translate\", \"value\": " + fmt3(cx, cy, cz) + "}]}";
  }
  Mesh ground = quad({-8, 0, 8}, {16, 0, 0}, {0, 0, -16}, {0, 1, 0});
  if (!writeObj(dir + "/ground.obj", ground)) return ASH_E_IO;
  {
    const float g = r.uf(0.15f, 0.8f);
    mats += ",\n    {\"name\": \"m_ground\", \"type\": \"brdf_lambertian\", \"diffuse_reflectance\": " +
            fmt3(g * r.uf(0.8f, 1.0f), g * r.uf(0.8f, 1.0f), g * r.uf(0.8f, 1.0f)) + "}";
  }
  meshes += ",\n    {\"name\": \"ground\", \"path\": \"ground.obj\"}";
  insts += ",\n    {\"mesh\": \"ground\", \"material\": \"m_ground\"}";
  // rect lights facing down: cross(v1 - p, v2 - p) points -y
  std::string lights;
  const int nLights = 1 + int(r.uni() * 4.0);
  for (int l = 0; l < nLights; ++l) {
    Rng q = objectRng(master, 3, 1000 + l);
    const double sx = q.uf(0.2f, 1.2f), sz = q.uf(0.2f, 1.2f), y = q.uf(2.0f, 4.0f);
    const double px = q.uf(-2.5f, 2.5f) - sx * 0.5, pz = q.uf(-2.5f, 2.5f) - sz * 0.5;
    const float e = q.uf(2.0f, 25.0f);
    lights += std::string(lights.empty() ? "" : ",\n") + "    {\"type\": \"rect\", \"radiance\": " +
              fmt3(e * q.uf(0.7f, 1.0f), e * q.uf(0.7f, 1.0f), e * q.uf(0.7f, 1.0f)) + ", \"position\": " +
              fmt3(px, y, pz) + ", \"v1\": " + fmt3(px + sx, y, pz) + ", \"v2\": " + fmt3(px, y, pz + sz) + "}";
  }
  // env map 256 x 128
  const int envW = 256, envH = 128;
  std::vector<float> env(size_t(envW) * envH * 4);
  const bool constant = r.uni() < 0.25;
  const float scale = r.uf(0.05f, 1.5f);
  const float cr = r.uf(0.3f, 1.0f), cg = r.uf(0.3f, 1.0f), cb = r.uf(0.3f, 1.0f);
  const double el = r.uf(5.0f, 70.0f) * M_PI / 180, az = r.uf(0.0f, 360.0f) * M_PI / 180;
  V3 sun = nrm(V3{std::cos(el) * std::cos(az), std::sin(el), std::cos(el) * std::sin(az)});
  const float sunL = r.uni() < 0.5 ? r.uf(50.0f, 2000.0f) : 0.0f;
  for (int j = 0; j < envH; ++j)
    for (int i = 0; i < envW; ++i) {
      double theta = (j + 0.5) / envH * M_PI, phi = (i + 0.5) / envW * 2 * M_PI - M_PI;
      V3 d{std::sin(theta) * std::cos(phi), std::cos(theta), std::sin(theta) * std::sin(phi)};
      float* px = &env[(size_t(j) * envW + i) * 4];
      if (constant) {
        px[0] = cr * scale, px[1] = cg * scale, px[2] = cb * scale;
      } else {
        skyColor(d, sun, px);
        px[0] *= cr * scale, px[1] *= cg * scale, px[2] *= cb * scale;
        if (dot(d, sun) > std::cos(2.0 * M_PI / 180.0)) px[0] = px[1] = px[2] = sunL;
      }
      px[3] = 1.f;
    }
  std::string err;
  if (!ash::writeImage(dir + "/sky.exr", uint32_t(envW), uint32_t(envH), env.data(), err)) return ASH_E_IO;
  std::string js = "{\n";
  js += "  \"state\": {\"path_tracing\": {\"spp\": 4, \"max_path_depth\": 6, "
        "\"multi_channel\": [\"diffuse\", \"normal\", \"depth\"]}, "
        "\"post_processing\": {\"tone_mapping\": \"filmic\"}, \"output_hdr\": true},\n";
  js += "  \"camera\": {\"type\": \"perspective\", \"film\": {\"resolution\": [256, 256]}, \"fov\": 50.0},\n";
  js += "  \"materials\": [\n" + mats + "\n  ],\n";
  js += "  \"lights\": [\n" + lights + "\n  ],\n";
  js += "  \"envmap\": {\"path\": \"sky.exr\"},\n";
  js += "  \"meshes\": [\n" + meshes + "\n  ],\n";
  js += "  \"instances\": [\n" + insts + "\n  ],\n";
  js += "  \"shots\": [\n    {\"type\": \"lookat\", \"eye\": [0, 2.5, 6], \"lookat\": [0, 0.4, 0], \"up\": [0, 1, 0]}\n  ]\n}\n";
  scenePath = dir + "/training.json";
  return writeText(scenePath, js) ? ASH_OK : ASH_E_IO;
}

}  // namespace

extern "C" int ash_generate_scene(int32_t config, uint64_t seed, int32_t scale, const char* dir, char* pathOut,
                                  size_t pathLen) {
  if (!dir) return ASH_E_INVALID;
  uint64_t master = seed ? seed : 0xA5A52026ull;
  std::string path;
  int rc;
  if (config == 0 || config == 2) {
    rc = genBox(master, config, dir, path);
  } else if (config == 1) {
    rc = scale > 0 ? genSpheres(master, 2, 3, 256, 320, 180, dir, path) : genSpheres(master, 8, 6, 2048, 1920, 1080, dir, path);
  } else if (config == 3) {
    rc = genTraining(master, dir, path);
  } else if (config == 4) {
    // resident serving scene of about `scale` triangles: the (grid, level)
    // whose grid^2 * 20 * 4^level is closest to it, grid in 2..8, level 2..6
    if (scale <= 0) return ASH_E_INVALID;
    int bg = 2, bl = 2;
    double best = 1e300;
    for (int g = 2; g <= 8; ++g)
      for (int l = 2; l <= 6; ++l) {
        double t = double(g) * g * 20.0 * std::pow(4.0, l);
        double e = std::fabs(std::log(t / double(scale)));
        if (e < best) best = e, bg = g, bl = l;
      }
    rc = genSpheres(master, bg, bl, 1024, 512, 512, dir, path);
  } else {
    return ASH_E_INVALID;
  }
  if (rc == ASH_OK && pathOut && pathLen) std::snprintf(pathOut, pathLen, "%s", path.c_str());
  return rc;
}
