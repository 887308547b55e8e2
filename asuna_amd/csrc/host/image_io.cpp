// Image IO for scene assets and render outputs, written from the file-format
// specifications (Radiance RGBE, OpenEXR scanline files, PNG, NumPy .npy):
//  - readImage  (src/core/texture.cpp:306-338): .hdr (RGBE -> float as
//    mantissa * 2^(e-136)), .exr (HALF/FLOAT channels; NONE/RLE/ZIPS/ZIP),
//    .png (8/16-bit; channels v/255 raised to `gamma`, alpha linear),
//    .jpg (jpeg.cpp; channels v/255 raised to `gamma`, alpha 1).
//  - writeImage (src/core/texture.cpp:340-372): .exr as HALF R,G,B with ZIP
//    compression (RgbaOutputFile WRITE_RGB), .hdr RGBE, .png 8-bit RGBA with
//    v = int(clamp(x*255 + 0.5)) (gamma 1), .npy float32.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "scene.h"

namespace ash {

namespace {

std::string extOf(const std::string& p) {
  size_t d = p.find_last_of('.');
  size_t s = p.find_last_of('/');
  if (d == std::string::npos || (s != std::string::npos && d < s)) return "";
  std::string e = p.substr(d + 1);
  for (auto& c : e) c = char(std::tolower(c));
  return e;
}

bool readFile(const std::string& p, std::vector<unsigned char>& out) {
  std::ifstream f(p, std::ios::binary);
  if (!f) return false;
  out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return true;
}

// ------------------------------------------------------------------ half
float halfToFloat(uint16_t h) {
  uint32_t s = uint32_t(h >> 15) << 31, e = (h >> 10) & 31, m = h & 1023, o;
  if (e == 0) {
    if (m == 0) {
      o = s;
    } else {  // subnormal
      int sh = 0;
      while (!(m & 1024)) m <<= 1, ++sh;
      m &= 1023;
      o = s | (uint32_t(127 - 15 - sh + 1) << 23) | (m << 13);
    }
  } else if (e == 31) {
    o = s | 0x7f800000u | (m << 13);
  } else {
    o = s | ((e + 112) << 23) | (m << 13);
  }
  float f;
  std::memcpy(&f, &o, 4);
  return f;
}

uint16_t floatToHalf(float f) {  // round to nearest even
  uint32_t x;
  std::memcpy(&x, &f, 4);
  uint32_t s = (x >> 16) & 0x8000u;
  uint32_t e = (x >> 23) & 0xff;
  uint32_t m = x & 0x7fffffu;
  if (e == 0xff) return uint16_t(s | 0x7c00u | (m ? 0x200u : 0u));
  int ne = int(e) - 127 + 15;
  if (ne >= 31) return uint16_t(s | 0x7c00u);
  if (ne <= 0) {
    if (ne < -10) return uint16_t(s);
    m |= 0x800000u;
    int shift = 14 - ne;
    uint32_t hm = m >> shift;
    uint32_t rem = m & ((1u << shift) - 1), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (hm & 1))) ++hm;
    return uint16_t(s | hm);
  }
  uint32_t hm = m >> 13, rem = m & 0x1fffu;
  uint32_t out = s | (uint32_t(ne) << 10) | hm;
  if (rem > 0x1000u || (rem == 0x1000u && (hm & 1))) ++out;  // may carry into exponent
  return uint16_t(out);
}

// ------------------------------------------------------------- RGBE (.hdr)
bool readHdr(const std::vector<unsigned char>& b, Image& im, std::string& err) {
  size_t p = 0;
  auto line = [&]() {
    std::string l;
    while (p < b.size() && b[p] != '\n') l += char(b[p++]);
    if (p < b.size()) ++p;
    return l;
  };
  std::string magic = line();
  if (magic != "#?RADIANCE" && magic != "#?RGBE") {
    err = "hdr: bad magic";
    return false;
  }
  bool fmt = false;
  for (;;) {
    if (p >= b.size()) {
      err = "hdr: truncated header";
      return false;
    }
    std::string l = line();
    if (l.empty()) break;
    if (l == "FORMAT=32-bit_rle_rgbe") fmt = true;
  }
  if (!fmt) {
    err = "hdr: unsupported format";
    return false;
  }
  std::string res = line();
  int h = 0, w = 0;
  if (std::sscanf(res.c_str(), "-Y %d +X %d", &h, &w) != 2 || w <= 0 || h <= 0) {
    err = "hdr: unsupported resolution line";
    return false;
  }
  im.w = uint32_t(w);
  im.h = uint32_t(h);
  im.rgba.assign(size_t(w) * h * 4, 0.f);
  std::vector<unsigned char> scan(size_t(w) * 4);
  auto conv = [&](size_t pix, const unsigned char* e) {
    float* o = &im.rgba[pix * 4];
    if (e[3]) {
      float f1 = std::ldexp(1.0f, int(e[3]) - int(128 + 8));
      o[0] = e[0] * f1, o[1] = e[1] * f1, o[2] = e[2] * f1;
    } else {
      o[0] = o[1] = o[2] = 0.f;
    }
    o[3] = 1.f;
  };
  for (int y = 0; y < h; ++y) {
    bool rle = false;
    if (w >= 8 && w < 32768 && p + 4 <= b.size() && b[p] == 2 && b[p + 1] == 2 && !(b[p + 2] & 0x80)) {
      int len = (b[p + 2] << 8) | b[p + 3];
      if (len != w) {
        err = "hdr: bad scanline length";
        return false;
      }
      rle = true;
      p += 4;
    }
    if (rle) {
      for (int c = 0; c < 4; ++c) {
        int x = 0;
        while (x < w) {
          if (p >= b.size()) {
            err = "hdr: truncated";
            return false;
          }
          int count = b[p++];
          if (count > 128) {
            count -= 128;
            if (p >= b.size() || x + count > w) {
              err = "hdr: bad run";
              return false;
            }
            unsigned char v = b[p++];
            for (int i = 0; i < count; ++i) scan[size_t(x++) * 4 + c] = v;
          } else {
            if (count == 0 || p + count > b.size() || x + count > w) {
              err = "hdr: bad dump";
              return false;
            }
            for (int i = 0; i < count; ++i) scan[size_t(x++) * 4 + c] = b[p++];
          }
        }
      }
    } else {
      if (p + size_t(w) * 4 > b.size()) {
        err = "hdr: truncated flat scanline";
        return false;
      }
      std::memcpy(scan.data(), &b[p], size_t(w) * 4);
      p += size_t(w) * 4;
    }
    for (int x = 0; x < w; ++x) conv(size_t(y) * w + x, &scan[size_t(x) * 4]);
  }
  return true;
}

bool writeHdr(const std::string& path, uint32_t w, uint32_t h, const float* rgba, std::string& err) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  std::fprintf(f, "#?RADIANCE\n# Written by asuna\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=          1.0000000000000\n\n-Y %u +X %u\n", h, w);
  std::vector<unsigned char> scan(size_t(w) * 4);
  for (uint32_t y = 0; y < h; ++y) {
    for (uint32_t x = 0; x < w; ++x) {
      const float* c = rgba + (size_t(y) * w + x) * 4;
      float mx = std::max(c[0], std::max(c[1], c[2]));
      unsigned char* e = &scan[size_t(x) * 4];
      if (mx < 1e-32f) {
        e[0] = e[1] = e[2] = e[3] = 0;
      } else {
        int ex;
        float norm = float(std::frexp(mx, &ex)) * 256.0f / mx;
        e[0] = (unsigned char)(c[0] * norm);
        e[1] = (unsigned char)(c[1] * norm);
        e[2] = (unsigned char)(c[2] * norm);
        e[3] = (unsigned char)(ex + 128);
      }
    }
    if (w < 8 || w >= 32768) {
      std::fwrite(scan.data(), 1, scan.size(), f);
      continue;
    }
    unsigned char hdr[4] = {2, 2, (unsigned char)(w >> 8), (unsigned char)(w & 255)};
    std::fwrite(hdr, 1, 4, f);
    for (int c = 0; c < 4; ++c) {
      uint32_t x = 0;
      while (x < w) {  // literal dumps of up to 128 bytes
        uint32_t n = std::min<uint32_t>(128, w - x);
        std::fputc(int(n), f);
        for (uint32_t i = 0; i < n; ++i) std::fputc(scan[size_t(x + i) * 4 + c], f);
        x += n;
      }
    }
  }
  std::fclose(f);
  return true;
}

// ------------------------------------------------------------------- EXR
struct ExrChan {
  std::string name;
  int type;  // 0 uint, 1 half, 2 float
};

void exrUnpredict(std::vector<unsigned char>& t) {
  for (size_t i = 1; i < t.size(); ++i) t[i] = (unsigned char)(int(t[i - 1]) + int(t[i]) - 128);
  std::vector<unsigned char> o(t.size());
  size_t half = (t.size() + 1) / 2;
  for (size_t i = 0; i < t.size(); ++i) o[i] = (i & 1) ? t[half + i / 2] : t[i / 2];
  t.swap(o);
}

void exrPredict(std::vector<unsigned char>& raw) {
  std::vector<unsigned char> t(raw.size());
  size_t half = (raw.size() + 1) / 2;
  for (size_t i = 0; i < raw.size(); ++i) {
    if (i & 1) t[half + i / 2] = raw[i];
    else t[i / 2] = raw[i];
  }
  for (size_t i = t.size(); i-- > 1;) t[i] = (unsigned char)(int(t[i]) - int(t[i - 1]) + 128);
  raw.swap(t);
}

bool readExr(const std::vector<unsigned char>& b, Image& im, std::string& err) {
  if (b.size() < 8 || b[0] != 0x76 || b[1] != 0x2f || b[2] != 0x31 || b[3] != 0x01) {
    err = "exr: bad magic";
    return false;
  }
  if (b[5] & 0x16) {  // tiled / long names / deep / multipart bits
    if (b[5] & 0x02) {
      err = "exr: tiled files are not supported";
      return false;
    }
  }
  size_t p = 8;
  auto cstr = [&]() {
    std::string s;
    while (p < b.size() && b[p]) s += char(b[p++]);
    ++p;
    return s;
  };
  auto rd32 = [&](size_t at) {
    int32_t v;
    std::memcpy(&v, &b[at], 4);
    return v;
  };
  std::vector<ExrChan> chans;
  int comp = -1;
  int32_t dw[4] = {0, 0, -1, -1};
  for (;;) {
    if (p >= b.size()) {
      err = "exr: truncated header";
      return false;
    }
    std::string name = cstr();
    if (name.empty()) break;
    std::string type = cstr();
    if (p + 4 > b.size()) {
      err = "exr: truncated attribute";
      return false;
    }
    int32_t sz = rd32(p);
    p += 4;
    size_t vs = p;
    if (vs + size_t(sz) > b.size()) {
      err = "exr: truncated attribute value";
      return false;
    }
    if (name == "channels") {
      size_t q = vs;
      while (q < vs + size_t(sz) && b[q]) {
        std::string cn;
        while (b[q]) cn += char(b[q++]);
        ++q;
        int32_t t;
        std::memcpy(&t, &b[q], 4);
        q += 16;
        chans.push_back({cn, t});
      }
    } else if (name == "compression") {
      comp = b[vs];
    } else if (name == "dataWindow") {
      for (int k = 0; k < 4; ++k) dw[k] = rd32(vs + 4 * k);
    }
    p = vs + size_t(sz);
  }
  int w = dw[2] - dw[0] + 1, h = dw[3] - dw[1] + 1;
  if (w <= 0 || h <= 0) {
    err = "exr: bad data window";
    return false;
  }
  int lpb;
  switch (comp) {
    case 0: case 1: case 2: lpb = 1; break;  // NONE, RLE, ZIPS
    case 3: lpb = 16; break;                 // ZIP
    default: err = "exr: unsupported compression " + std::to_string(comp); return false;
  }
  size_t rowBytes = 0;
  for (auto& c : chans) rowBytes += size_t(w) * (c.type == 1 ? 2 : 4);
  int nChunks = (h + lpb - 1) / lpb;
  std::vector<uint64_t> offs(static_cast<size_t>(nChunks));
  if (p + size_t(nChunks) * 8 > b.size()) {
    err = "exr: truncated offset table";
    return false;
  }
  std::memcpy(offs.data(), &b[p], size_t(nChunks) * 8);
  im.w = uint32_t(w);
  im.h = uint32_t(h);
  im.rgba.assign(size_t(w) * h * 4, 0.f);
  for (size_t i = 0; i < size_t(w) * h; ++i) im.rgba[i * 4 + 3] = 1.f;
  bool hasA = false;
  for (auto& c : chans) hasA |= c.name == "A";
  (void)hasA;
  for (int ci = 0; ci < nChunks; ++ci) {
    size_t q = size_t(offs[size_t(ci)]);
    if (q + 8 > b.size()) {
      err = "exr: bad chunk offset";
      return false;
    }
    int32_t y0 = rd32(q), dsz = rd32(q + 4);
    q += 8;
    if (q + size_t(dsz) > b.size()) {
      err = "exr: truncated chunk";
      return false;
    }
    int rows = std::min(lpb, dw[3] + 1 - y0);
    size_t expect = rowBytes * size_t(rows);
    std::vector<unsigned char> data;
    if (size_t(dsz) == expect) {
      data.assign(&b[q], &b[q] + expect);  // stored uncompressed
    } else if (comp == 2 || comp == 3) {
      data.resize(expect);
      uLongf dl = uLongf(expect);
      if (uncompress(data.data(), &dl, &b[q], uLong(dsz)) != Z_OK || dl != expect) {
        err = "exr: zlib error";
        return false;
      }
      exrUnpredict(data);
    } else if (comp == 1) {
      std::vector<unsigned char> t;
      size_t r = q, e = q + size_t(dsz);
      while (r < e) {
        int n = int(int8_t(b[r++]));
        if (n < 0) {
          for (int i = 0; i < -n && r < e; ++i) t.push_back(b[r++]);
        } else if (r < e) {
          unsigned char v = b[r++];
          for (int i = 0; i <= n; ++i) t.push_back(v);
        }
      }
      if (t.size() != expect) {
        err = "exr: bad RLE chunk";
        return false;
      }
      exrUnpredict(t);
      data.swap(t);
    } else {
      err = "exr: unexpected chunk size";
      return false;
    }
    size_t r = 0;
    for (int yy = 0; yy < rows; ++yy) {
      int y = y0 + yy - dw[1];
      for (auto& c : chans) {
        int slot = c.name == "R" ? 0 : c.name == "G" ? 1 : c.name == "B" ? 2 : c.name == "A" ? 3 : -1;
        for (int x = 0; x < w; ++x) {
          float v;
          if (c.type == 1) {
            uint16_t hv;
            std::memcpy(&hv, &data[r], 2);
            r += 2;
            v = halfToFloat(hv);
          } else if (c.type == 2) {
            std::memcpy(&v, &data[r], 4);
            r += 4;
          } else {
            uint32_t u;
            std::memcpy(&u, &data[r], 4);
            r += 4;
            v = float(u);
          }
          if (slot >= 0 && y >= 0 && y < h) im.rgba[(size_t(y) * w + x) * 4 + size_t(slot)] = v;
        }
      }
    }
  }
  return true;
}

void put32(std::vector<unsigned char>& o, int32_t v) {
  unsigned char b[4];
  std::memcpy(b, &v, 4);
  o.insert(o.end(), b, b + 4);
}
void putStr(std::vector<unsigned char>& o, const char* s) { o.insert(o.end(), s, s + std::strlen(s) + 1); }

bool writeExr(const std::string& path, uint32_t w, uint32_t h, const float* rgba, std::string& err) {
  // Scanline file, HALF channels B,G,R (alphabetical), ZIP (16 lines/chunk).
  std::vector<unsigned char> o = {0x76, 0x2f, 0x31, 0x01, 2, 0, 0, 0};
  putStr(o, "channels");
  putStr(o, "chlist");
  put32(o, 3 * 18 + 1);
  for (const char* c : {"B", "G", "R"}) {
    putStr(o, c);
    put32(o, 1);
    o.push_back(0), o.push_back(0), o.push_back(0), o.push_back(0);
    put32(o, 1);
    put32(o, 1);
  }
  o.push_back(0);
  putStr(o, "compression");
  putStr(o, "compression");
  put32(o, 1);
  o.push_back(3);
  for (const char* win : {"dataWindow", "displayWindow"}) {
    putStr(o, win);
    putStr(o, "box2i");
    put32(o, 16);
    put32(o, 0), put32(o, 0), put32(o, int32_t(w) - 1), put32(o, int32_t(h) - 1);
  }
  putStr(o, "lineOrder");
  putStr(o, "lineOrder");
  put32(o, 1);
  o.push_back(0);
  float one = 1.f, zero = 0.f;
  auto putF = [&](float f) {
    unsigned char bb[4];
    std::memcpy(bb, &f, 4);
    o.insert(o.end(), bb, bb + 4);
  };
  putStr(o, "pixelAspectRatio");
  putStr(o, "float");
  put32(o, 4);
  putF(one);
  putStr(o, "screenWindowCenter");
  putStr(o, "v2f");
  put32(o, 8);
  putF(zero), putF(zero);
  putStr(o, "screenWindowWidth");
  putStr(o, "float");
  put32(o, 4);
  putF(one);
  o.push_back(0);
  const uint32_t lpb = 16;
  uint32_t nChunks = (h + lpb - 1) / lpb;
  size_t tablePos = o.size();
  o.resize(o.size() + size_t(nChunks) * 8);
  std::vector<unsigned char> raw, comp;
  for (uint32_t ci = 0; ci < nChunks; ++ci) {
    uint32_t y0 = ci * lpb, rows = std::min(lpb, h - y0);
    raw.clear();
    for (uint32_t yy = 0; yy < rows; ++yy) {
      for (int c : {2, 1, 0}) {
        for (uint32_t x = 0; x < w; ++x) {
          uint16_t hv = floatToHalf(rgba[(size_t(y0 + yy) * w + x) * 4 + size_t(c)]);
          raw.push_back((unsigned char)(hv & 255));
          raw.push_back((unsigned char)(hv >> 8));
        }
      }
    }
    std::vector<unsigned char> pred = raw;
    exrPredict(pred);
    uLongf cl = compressBound(uLong(pred.size()));
    comp.resize(cl);
    if (compress2(comp.data(), &cl, pred.data(), uLong(pred.size()), Z_DEFAULT_COMPRESSION) != Z_OK) {
      err = "exr: zlib compress failed";
      return false;
    }
    uint64_t off = o.size();
    std::memcpy(&o[tablePos + size_t(ci) * 8], &off, 8);
    put32(o, int32_t(y0));
    if (cl < raw.size()) {
      put32(o, int32_t(cl));
      o.insert(o.end(), comp.begin(), comp.begin() + long(cl));
    } else {
      put32(o, int32_t(raw.size()));
      o.insert(o.end(), raw.begin(), raw.end());
    }
  }
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size();
  std::fclose(f);
  if (!ok) err = "write failed: " + path;
  return ok;
}

// ------------------------------------------------------------------- PNG
uint32_t be32(const unsigned char* p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]; }

bool readPng(const std::vector<unsigned char>& b, float gamma, Image& im, std::string& err) {
  static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (b.size() < 8 || std::memcmp(b.data(), sig, 8) != 0) {
    err = "png: bad signature";
    return false;
  }
  size_t p = 8;
  uint32_t w = 0, h = 0;
  int depth = 0, ctype = 0, interlace = 0;
  std::vector<unsigned char> idat, plte, trns;
  while (p + 8 <= b.size()) {
    uint32_t len = be32(&b[p]);
    std::string type(reinterpret_cast<const char*>(&b[p + 4]), 4);
    if (p + 12 + len > b.size()) {
      err = "png: truncated chunk";
      return false;
    }
    const unsigned char* d = &b[p + 8];
    if (type == "IHDR") {
      w = be32(d);
      h = be32(d + 4);
      depth = d[8];
      ctype = d[9];
      interlace = d[12];
    } else if (type == "PLTE") {
      plte.assign(d, d + len);
    } else if (type == "tRNS") {
      trns.assign(d, d + len);
    } else if (type == "IDAT") {
      idat.insert(idat.end(), d, d + len);
    } else if (type == "IEND") {
      break;
    }
    p += 12 + len;
  }
  if (!w || !h || interlace || (depth != 8 && depth != 16) || !(ctype == 0 || ctype == 2 || ctype == 3 || ctype == 4 || ctype == 6)) {
    err = "png: unsupported layout (need 8/16-bit, non-interlaced)";
    return false;
  }
  int ch = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4;
  size_t bpp = size_t(ch) * (depth / 8), stride = size_t(w) * bpp;
  std::vector<unsigned char> raw((stride + 1) * h);
  uLongf rl = uLongf(raw.size());
  if (uncompress(raw.data(), &rl, idat.data(), uLong(idat.size())) != Z_OK || rl != raw.size()) {
    err = "png: inflate failed";
    return false;
  }
  std::vector<unsigned char> img(stride * h), prev(stride, 0);
  for (uint32_t y = 0; y < h; ++y) {
    unsigned char f = raw[y * (stride + 1)];
    const unsigned char* s = &raw[y * (stride + 1) + 1];
    unsigned char* o = &img[y * stride];
    for (size_t i = 0; i < stride; ++i) {
      int a = i >= bpp ? o[i - bpp] : 0, up = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
      int v = s[i];
      switch (f) {
        case 0: break;
        case 1: v += a; break;
        case 2: v += up; break;
        case 3: v += (a + up) / 2; break;
        case 4: {
          int pp = a + up - c, pa = std::abs(pp - a), pb = std::abs(pp - up), pc = std::abs(pp - c);
          v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? up : c);
          break;
        }
        default: err = "png: bad filter"; return false;
      }
      o[i] = (unsigned char)v;
    }
    std::memcpy(prev.data(), o, stride);
  }
  im.w = w;
  im.h = h;
  im.rgba.assign(size_t(w) * h * 4, 0.f);
  auto toHdr = [&](int v) { return float(std::pow(double(float(v) / 255.0f), double(gamma))); };
  for (size_t i = 0; i < size_t(w) * h; ++i) {
    int c8[4] = {0, 0, 0, 255};
    auto sample = [&](int k) { return depth == 8 ? img[i * bpp + size_t(k)] : img[i * bpp + size_t(k) * 2]; };
    if (ctype == 0) c8[0] = c8[1] = c8[2] = sample(0);
    else if (ctype == 2) c8[0] = sample(0), c8[1] = sample(1), c8[2] = sample(2);
    else if (ctype == 3) {
      int idx = sample(0);
      if (size_t(idx) * 3 + 2 < plte.size()) c8[0] = plte[idx * 3], c8[1] = plte[idx * 3 + 1], c8[2] = plte[idx * 3 + 2];
      if (size_t(idx) < trns.size()) c8[3] = trns[size_t(idx)];
    } else if (ctype == 4) c8[0] = c8[1] = c8[2] = sample(0), c8[3] = sample(1);
    else c8[0] = sample(0), c8[1] = sample(1), c8[2] = sample(2), c8[3] = sample(3);
    for (int k = 0; k < 3; ++k) im.rgba[i * 4 + size_t(k)] = toHdr(c8[k]);
    im.rgba[i * 4 + 3] = float(c8[3]) / 255.0f;
  }
  return true;
}

void pngChunk(FILE* f, const char* type, const unsigned char* d, uint32_t len) {
  unsigned char hdr[8] = {(unsigned char)(len >> 24), (unsigned char)(len >> 16), (unsigned char)(len >> 8),
                          (unsigned char)len, (unsigned char)type[0], (unsigned char)type[1],
                          (unsigned char)type[2], (unsigned char)type[3]};
  std::fwrite(hdr, 1, 8, f);
  if (len) std::fwrite(d, 1, len, f);
  uLong crc = crc32(0, hdr + 4, 4);
  if (len) crc = crc32(crc, d, len);
  unsigned char c[4] = {(unsigned char)(crc >> 24), (unsigned char)(crc >> 16), (unsigned char)(crc >> 8), (unsigned char)crc};
  std::fwrite(c, 1, 4, f);
}

bool writePng(const std::string& path, uint32_t w, uint32_t h, const float* rgba, std::string& err) {
  std::vector<unsigned char> raw((size_t(w) * 4 + 1) * h);
  for (uint32_t y = 0; y < h; ++y) {
    unsigned char* o = &raw[y * (size_t(w) * 4 + 1)];
    *o++ = 0;
    for (uint32_t x = 0; x < size_t(w) * 4; ++x) {
      float v = rgba[size_t(y) * w * 4 + x] * 255.f + 0.5f;  // stbi__hdr_to_ldr with gamma 1
      v = std::min(255.f, std::max(0.f, v));
      *o++ = (unsigned char)int(v);
    }
  }
  uLongf cl = compressBound(uLong(raw.size()));
  std::vector<unsigned char> comp(cl);
  if (compress2(comp.data(), &cl, raw.data(), uLong(raw.size()), 6) != Z_OK) {
    err = "png: deflate failed";
    return false;
  }
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  std::fwrite(sig, 1, 8, f);
  unsigned char ihdr[13] = {(unsigned char)(w >> 24), (unsigned char)(w >> 16), (unsigned char)(w >> 8), (unsigned char)w,
                            (unsigned char)(h >> 24), (unsigned char)(h >> 16), (unsigned char)(h >> 8), (unsigned char)h,
                            8, 6, 0, 0, 0};
  pngChunk(f, "IHDR", ihdr, 13);
  pngChunk(f, "IDAT", comp.data(), uint32_t(cl));
  pngChunk(f, "IEND", nullptr, 0);
  std::fclose(f);
  return true;
}

}  // namespace

bool readImage(const std::string& path, float gamma, Image& out, std::string& err) {
  std::string e = extOf(path);
  if (e != "hdr" && e != "exr" && e != "jpg" && e != "png") {
    err = "Scene: textures only support extensions (hdr exr jpg png) while [" + e + "] is passed in";
    return false;
  }
  std::vector<unsigned char> b;
  if (!readFile(path, b)) {
    err = "Scene Error: failed to load " + path;
    return false;
  }
  bool ok;
  if (e == "hdr") ok = readHdr(b, out, err);
  else if (e == "exr") ok = readExr(b, out, err);
  else if (e == "png") ok = readPng(b, gamma, out, err);
  else ok = readJpeg(b, gamma, out, err);
  if (!ok) err = "Scene Error: failed to read image file " + path + ": " + err;
  return ok;
}

bool writeImage(const std::string& path, uint32_t w, uint32_t h, const float* rgba, std::string& err) {
  std::string e = extOf(path);
  if (e == "hdr") return writeHdr(path, w, h, rgba, err);
  if (e == "exr") return writeExr(path, w, h, rgba, err);
  if (e == "png") return writePng(path, w, h, rgba, err);
  if (e == "npy") return writeNpy(path, {h, w, 4}, rgba, err);
  err = "Scene Error: output supports extensions (hdr exr png npy) while [" + e + "] is passed in";
  return false;
}

bool writeNpy(const std::string& path, const std::vector<unsigned long>& shape, const float* data, std::string& err) {
  std::ostringstream dict;
  dict << "{'descr': '<f4', 'fortran_order': False, 'shape': (";
  size_t n = 1;
  for (size_t i = 0; i < shape.size(); ++i) {
    dict << shape[i] << (shape.size() == 1 ? "," : (i + 1 < shape.size() ? ", " : ""));
    n *= shape[i];
  }
  dict << "), }";
  std::string hs = dict.str();
  size_t total = 10 + hs.size() + 1;
  hs.append((64 - total % 64) % 64, ' ');
  hs += '\n';
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  std::fwrite(magic, 1, 8, f);
  uint16_t hl = uint16_t(hs.size());
  std::fwrite(&hl, 2, 1, f);
  std::fwrite(hs.data(), 1, hs.size(), f);
  std::fwrite(data, 4, n, f);
  std::fclose(f);
  return true;
}

// Uncharted 2 curve, utils/tonemapping.glsl:43-54
static float uncharted2(float x) {
  const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
  return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}

// pcg3d integer hash, utils/random.glsl:76-86
static void pcg3d(uint32_t v[3]) {
  for (int k = 0; k < 3; ++k) v[k] = v[k] * 1664525u + 1013904223u;
  v[0] += v[1] * v[2];
  v[1] += v[2] * v[0];
  v[2] += v[0] * v[1];
  for (int k = 0; k < 3; ++k) v[k] ^= v[k] >> 16;
  v[0] += v[1] * v[2];
  v[1] += v[2] * v[0];
  v[2] += v[0] * v[1];
}

// The custom operator with the post state the loader leaves at its
// defaults (state.h:47-58; loader.cpp:199-227 reads only the type):
// exposure avgLum = 1, Uncharted 2 + gamma (TONEMAP_UNCHARTED), dither to
// 1/255 with pcg3d noise of the pixel, contrast / brightness / saturation 1,
// vignette 0, no auto exposure.  post.idle.frag:108-131.
static void toneMapCustom(uint32_t x, uint32_t y, float* c) {
  const float g = 2.2f, ig = 1.0f / 2.2f, quant = 1.0f / 255.0f;
  const float white = 1.0f / uncharted2(11.2f);
  uint32_t r[3] = {x, y, 0u};
  pcg3d(r);
  for (int k = 0; k < 3; ++k) {
    float t = std::pow(uncharted2(c[k] * 1.0f * 2.0f) * white, ig);  // toneMap(hdr, avgLum)
    float lin = std::pow(t, g);                                       // sRGBToLinear
    uint32_t bits = 0x3f800000u | (r[k] >> 9);
    float noise;
    std::memcpy(&noise, &bits, 4);
    noise -= 1.0f;
    float c0 = std::floor(std::pow(lin, ig) / quant) * quant;  // dither()
    float c1 = c0 + quant;
    float discr = std::pow(c0, g) * (1.0f - noise) + std::pow(c1, g) * noise;
    float v = discr < lin ? c1 : c0;
    v = std::min(1.0f, std::max(0.0f, v));  // contrast 1 then clamp; brightness, saturation 1; vignette 0
    c[k] = v;
  }
}

// post.idle.frag:71-134 tone mapping operators; pixel (x, y) = index % w, / w.
void toneMap(uint32_t type, uint32_t w, uint32_t h, float* rgba) {
  const float invGamma = 1.0f / 2.2f;
  const size_t n = size_t(w) * h;
  for (size_t i = 0; i < n; ++i) {
    float* c = rgba + i * 4;
    if (type == 6) {
      toneMapCustom(uint32_t(i % w), uint32_t(i / w), c);
      continue;
    }
    for (int k = 0; k < 3; ++k) {
      float x = c[k];
      float y = x;
      switch (type) {
        case 0: y = x; break;
        case 1: y = std::pow(x / (1.f + x / 1.5f), invGamma); break;
        case 2:
        case 4: {
          float v = std::max(0.0f, x - 0.004f);
          y = (v * (6.2f * v + 0.5f)) / (v * (6.2f * v + 1.7f) + 0.06f);
          break;
        }
        case 3: {
          float v = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
          y = std::pow(std::min(1.f, std::max(0.f, v)), invGamma);
          break;
        }
        case 5: y = x < 0.0031308f ? 12.92f * x : 1.055f * std::pow(x, 1.0f / 2.4f) - 0.055f; break;
        default: y = std::pow(std::max(0.f, x), invGamma); break;
      }
      c[k] = y;
    }
  }
}

}  // namespace ash
