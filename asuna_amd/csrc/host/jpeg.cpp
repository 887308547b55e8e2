// JPEG decoder for .jpg textures (readImage, src/core/texture.cpp:306-327,
// where the reference reads jpg through stb_image), written from ITU-T T.81:
// Huffman-coded sequential (SOF0/SOF1) and progressive (SOF2) DCT images of
// 8-bit precision, one (grey) or three components with any sampling factors
// up to 4, restart intervals.  Samples are reconstructed with an exact
// separable inverse DCT, 2x chroma is upsampled with the triangle filter
// ("fancy" upsampling) and other ratios by replication, YCbCr goes to RGB
// with the JFIF matrix.  Arithmetic coding, 12-bit samples, lossless and
// four-component (CMYK) files are rejected with an error.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "scene.h"

namespace ash {

namespace {

// Zig-zag scan position -> natural (row-major) coefficient index.
constexpr uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Canonical Huffman table (T.81 Annex C): code lengths -> (min, max) code per
// length and the index of its first value.
struct Huffman {
  uint8_t counts[17] = {0};
  uint8_t values[256] = {0};
  int32_t minCode[17] = {0}, maxCode[18] = {0}, first[17] = {0};
  bool defined = false;
  void build() {
    int32_t code = 0, k = 0;
    for (int len = 1; len <= 16; ++len) {
      first[len] = k;
      minCode[len] = code;
      code += counts[len];
      k += counts[len];
      maxCode[len] = counts[len] ? code - 1 : -1;
      code <<= 1;
    }
    maxCode[17] = 0x7fffffff;
    defined = true;
  }
};

struct Component {
  int id = 0, h = 1, v = 1, tq = 0;
  int td = 0, ta = 0;  // DC / AC table of the current scan
  int bw = 0, bh = 0;  // blocks per row / column, padded to whole MCUs
  int dcPred = 0;
  std::vector<int16_t> coef;  // bw * bh blocks of 64 coefficients (natural order)
};

// Entropy-coded segment reader: byte stuffing (FF 00), markers end the data
// (zeros are fed after one).
struct Bits {
  const std::vector<unsigned char>& b;
  size_t pos;
  uint32_t acc = 0;
  int n = 0;
  bool atMarker = false;
  Bits(const std::vector<unsigned char>& buf, size_t p) : b(buf), pos(p) {}
  void fill() {
    while (n <= 24) {
      uint32_t byte = 0;
      if (!atMarker && pos < b.size()) {
        byte = b[pos];
        if (byte == 0xFF) {
          const uint32_t next = pos + 1 < b.size() ? b[pos + 1] : 0xD9;
          if (next == 0x00) pos += 2;
          else {
            atMarker = true;
            byte = 0;
          }
        } else {
          ++pos;
        }
      }
      acc |= byte << (24 - n);
      n += 8;
    }
  }
  int bit() {
    if (n < 1) fill();
    const int r = int(acc >> 31);
    acc <<= 1;
    --n;
    return r;
  }
  int get(int k) {
    if (k == 0) return 0;
    if (n < k) fill();
    const int r = int(acc >> (32 - k));
    acc <<= k;
    n -= k;
    return r;
  }
  // T.81 F.2.2.1 EXTEND: a k-bit magnitude category value
  int receive(int k) {
    if (k == 0) return 0;
    const int v = get(k);
    return v < (1 << (k - 1)) ? v - (1 << k) + 1 : v;
  }
  int decode(const Huffman& h) {
    int32_t code = 0;
    for (int len = 1; len <= 16; ++len) {
      code = (code << 1) | bit();
      if (code <= h.maxCode[len]) return h.values[h.first[len] + code - h.minCode[len]];
    }
    return -1;  // corrupt data
  }
  // At a restart marker: drop the partial byte and step over RSTn.
  void restart() {
    acc = 0;
    n = 0;
    atMarker = false;
    while (pos + 1 < b.size() && !(b[pos] == 0xFF && b[pos + 1] >= 0xD0 && b[pos + 1] <= 0xD7)) ++pos;
    if (pos + 1 < b.size()) pos += 2;
  }
};

// Exact 8x8 inverse DCT (T.81 A.3.3), level shift and clamp to 0..255.
void idct8x8(const int16_t* in, const uint16_t* q, uint8_t* out, int stride) {
  static double cosTab[8][8];
  static bool init = false;
  if (!init) {
    for (int x = 0; x < 8; ++x)
      for (int u = 0; u < 8; ++u)
        cosTab[x][u] = (u == 0 ? std::sqrt(0.5) : 1.0) * std::cos((2 * x + 1) * u * M_PI / 16.0);
    init = true;
  }
  double f[64], t[64];
  for (int i = 0; i < 64; ++i) f[i] = double(in[i]) * double(q[i]);
  for (int v = 0; v < 8; ++v)  // rows: t[v][x] = 1/2 sum_u C(u) F(v,u) cos
    for (int x = 0; x < 8; ++x) {
      double s = 0;
      for (int u = 0; u < 8; ++u) s += cosTab[x][u] * f[v * 8 + u];
      t[v * 8 + x] = 0.5 * s;
    }
  for (int y = 0; y < 8; ++y)
    for (int x = 0; x < 8; ++x) {
      double s = 0;
      for (int v = 0; v < 8; ++v) s += cosTab[y][v] * t[v * 8 + x];
      const long r = std::lround(0.5 * s + 128.0);
      out[y * stride + x] = uint8_t(std::min(255L, std::max(0L, r)));
    }
}

struct Decoder {
  const std::vector<unsigned char>& b;
  std::string& err;
  uint16_t qt[4][64] = {{0}};  // natural order
  Huffman dc[4], ac[4];
  std::vector<Component> comps;
  int width = 0, height = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
  int restartInterval = 0;
  bool progressive = false, frame = false, adobe = false, jfif = false;
  int adobeTransform = -1;
  int eobrun = 0;
  Decoder(const std::vector<unsigned char>& buf, std::string& e) : b(buf), err(e) {}

  bool fail(const char* m) {
    err = std::string("jpg: ") + m;
    return false;
  }
  int u16(size_t p) const { return (int(b[p]) << 8) | b[p + 1]; }

  bool readDqt(size_t p, size_t end) {
    while (p < end) {
      const int pq = b[p] >> 4, tq = b[p] & 15;
      if (tq > 3 || pq > 1) return fail("bad quantisation table");
      ++p;
      for (int i = 0; i < 64; ++i) {
        if (p + (pq ? 2 : 1) > end) return fail("truncated quantisation table");
        qt[tq][kZigzag[i]] = uint16_t(pq ? u16(p) : b[p]);
        p += pq ? 2 : 1;
      }
    }
    return true;
  }
  bool readDht(size_t p, size_t end) {
    while (p < end) {
      const int tc = b[p] >> 4, th = b[p] & 15;
      if (tc > 1 || th > 3 || p + 17 > end) return fail("bad Huffman table");
      Huffman& h = tc ? ac[th] : dc[th];
      int total = 0;
      for (int i = 1; i <= 16; ++i) total += (h.counts[i] = b[p + size_t(i)]);
      p += 17;
      if (total > 256 || p + size_t(total) > end) return fail("bad Huffman table");
      std::memcpy(h.values, &b[p], size_t(total));
      p += size_t(total);
      h.build();
    }
    return true;
  }
  bool readSof(size_t p, size_t end) {
    if (frame) return fail("more than one frame");
    if (end < p + 6 || b[p] != 8) return fail("only 8-bit samples are supported");
    height = u16(p + 1);
    width = u16(p + 3);
    const int nc = b[p + 5];
    if (!width || !height) return fail("zero image size (DNL is not supported)");
    if (nc != 1 && nc != 3) return fail("only 1- and 3-component images are supported");
    if (end < p + 6 + 3 * size_t(nc)) return fail("truncated frame header");
    comps.resize(size_t(nc));
    for (int i = 0; i < nc; ++i) {
      Component& c = comps[size_t(i)];
      c.id = b[p + 6 + 3 * size_t(i)];
      c.h = b[p + 7 + 3 * size_t(i)] >> 4;
      c.v = b[p + 7 + 3 * size_t(i)] & 15;
      c.tq = b[p + 8 + 3 * size_t(i)];
      if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) return fail("bad component parameters");
      hmax = std::max(hmax, c.h);
      vmax = std::max(vmax, c.v);
    }
    mcux = (width + 8 * hmax - 1) / (8 * hmax);
    mcuy = (height + 8 * vmax - 1) / (8 * vmax);
    for (Component& c : comps) {
      c.bw = mcux * c.h;
      c.bh = mcuy * c.v;
      c.coef.assign(size_t(c.bw) * size_t(c.bh) * 64, 0);
    }
    frame = true;
    return true;
  }

  // One block of a scan (T.81 F.2.2 sequential; G.1.2 progressive).
  bool block(Bits& bits, Component& c, int16_t* blk, int ss, int se, int ah, int al) {
    if (!progressive) {
      const int t = bits.decode(dc[c.td]);
      if (t < 0 || t > 11) return fail("corrupt DC code");
      c.dcPred += bits.receive(t);
      blk[0] = int16_t(c.dcPred);
      for (int k = 1; k < 64;) {
        const int rs = bits.decode(ac[c.ta]);
        if (rs < 0) return fail("corrupt AC code");
        const int r = rs >> 4, s = rs & 15;
        if (s == 0) {
          if (r != 15) break;
          k += 16;
          continue;
        }
        k += r;
        if (k > 63) return fail("AC coefficient index out of range");
        blk[kZigzag[k++]] = int16_t(bits.receive(s));
      }
      return true;
    }
    if (ss == 0) {  // DC scans
      if (ah == 0) {
        const int t = bits.decode(dc[c.td]);
        if (t < 0 || t > 11) return fail("corrupt DC code");
        c.dcPred += bits.receive(t);
        blk[0] = int16_t(c.dcPred * (1 << al));
      } else if (bits.bit()) {
        blk[0] = int16_t(blk[0] | (1 << al));
      }
      return true;
    }
    if (ah == 0) {  // AC first pass
      if (eobrun > 0) {
        --eobrun;
        return true;
      }
      for (int k = ss; k <= se; ++k) {
        const int rs = bits.decode(ac[c.ta]);
        if (rs < 0) return fail("corrupt AC code");
        const int r = rs >> 4, s = rs & 15;
        if (s == 0) {
          if (r < 15) {
            eobrun = (1 << r) - 1 + bits.get(r);
            break;
          }
          k += 15;
          continue;
        }
        k += r;
        if (k > 63) return fail("AC coefficient index out of range");
        blk[kZigzag[k]] = int16_t(bits.receive(s) * (1 << al));
      }
      return true;
    }
    // AC refinement (G.1.2.3): one correction bit for every coefficient
    // already non-zero, new coefficients of magnitude 1 << al placed after
    // the coded number of still-zero ones.
    const int p1 = 1 << al, m1 = -(1 << al);
    int k = ss;
    auto refine = [&](int16_t& v) {
      if (bits.bit() && (v & p1) == 0) v = int16_t(v >= 0 ? v + p1 : v + m1);
    };
    if (eobrun == 0) {
      for (; k <= se; ++k) {
        const int rs = bits.decode(ac[c.ta]);
        if (rs < 0) return fail("corrupt AC code");
        int r = rs >> 4, s = rs & 15;
        int val = 0;
        if (s) {
          val = bits.bit() ? p1 : m1;
        } else if (r != 15) {
          eobrun = (1 << r) + bits.get(r);
          break;
        }
        for (; k <= se; ++k) {
          int16_t& v = blk[kZigzag[k]];
          if (v != 0) refine(v);
          else if (r-- == 0) break;
        }
        if (val && k <= se) blk[kZigzag[k]] = int16_t(val);
      }
    }
    if (eobrun > 0) {
      for (; k <= se; ++k) {
        int16_t& v = blk[kZigzag[k]];
        if (v != 0) refine(v);
      }
      --eobrun;
    }
    return true;
  }

  // Reads a scan header and its entropy-coded data; returns the position
  // after the data.
  bool scan(size_t p, size_t hdrEnd, size_t& after) {
    const int ns = b[p];
    if (ns < 1 || ns > 4 || hdrEnd < p + 1 + 2 * size_t(ns) + 3) return fail("bad scan header");
    std::vector<Component*> sc;
    for (int i = 0; i < ns; ++i) {
      const int id = b[p + 1 + 2 * size_t(i)], t = b[p + 2 + 2 * size_t(i)];
      Component* cp = nullptr;
      for (Component& c : comps)
        if (c.id == id) cp = &c;
      if (!cp) return fail("scan names an unknown component");
      cp->td = t >> 4;
      cp->ta = t & 15;
      if (cp->td > 3 || cp->ta > 3) return fail("bad table selector");
      sc.push_back(cp);
    }
    const size_t q = p + 1 + 2 * size_t(ns);
    const int ss = b[q], se = b[q + 1], ah = b[q + 2] >> 4, al = b[q + 2] & 15;
    if (progressive) {
      if (ss > se || se > 63 || (ss == 0 && se != 0) || (ss > 0 && ns != 1) || al > 13)
        return fail("bad progressive scan parameters");
    }
    for (Component* c : sc) {
      const bool needDc = !progressive || (ss == 0 && ah == 0);
      const bool needAc = !progressive ? true : ss > 0;
      if ((needDc && !dc[c->td].defined) || (needAc && !ac[c->ta].defined)) return fail("missing Huffman table");
      c->dcPred = 0;
    }
    eobrun = 0;
    Bits bits(b, hdrEnd);
    int todo = restartInterval;
    auto restartPoint = [&]() {
      if (!restartInterval) return;
      if (--todo == 0) {
        bits.restart();
        for (Component* c : sc) c->dcPred = 0;
        eobrun = 0;
        todo = restartInterval;
      }
    };
    if (ns == 1) {
      // non-interleaved: the component's own blocks in raster order
      Component& c = *sc[0];
      const int cw = (width * c.h + hmax - 1) / hmax, ch = (height * c.v + vmax - 1) / vmax;
      const int nbx = (cw + 7) / 8, nby = (ch + 7) / 8;
      for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx) {
          if (!block(bits, c, &c.coef[(size_t(by) * size_t(c.bw) + size_t(bx)) * 64], ss, se, ah, al)) return false;
          restartPoint();
        }
    } else {
      for (int my = 0; my < mcuy; ++my)
        for (int mx = 0; mx < mcux; ++mx) {
          for (Component* c : sc)
            for (int v = 0; v < c->v; ++v)
              for (int h = 0; h < c->h; ++h) {
                const size_t bx = size_t(mx) * size_t(c->h) + size_t(h), by = size_t(my) * size_t(c->v) + size_t(v);
                if (!block(bits, *c, &c->coef[(by * size_t(c->bw) + bx) * 64], ss, se, ah, al)) return false;
              }
          restartPoint();
        }
    }
    // continue after the entropy-coded data: the next marker that is not a
    // stuffed byte or a restart marker
    size_t e = hdrEnd;
    while (e + 1 < b.size() && !(b[e] == 0xFF && b[e + 1] != 0x00 && !(b[e + 1] >= 0xD0 && b[e + 1] <= 0xD7))) ++e;
    after = e;
    return true;
  }

  bool run() {
    if (b.size() < 4 || b[0] != 0xFF || b[1] != 0xD8) return fail("not a JPEG file (no SOI)");
    size_t p = 2;
    bool done = false;
    while (!done) {
      while (p < b.size() && b[p] != 0xFF) ++p;  // tolerate fill bytes
      while (p < b.size() && b[p] == 0xFF) ++p;
      if (p >= b.size()) return fail("missing EOI");
      const int m = b[p++];
      if (m == 0xD9) break;
      if (m >= 0xD0 && m <= 0xD7) continue;  // stray restart marker
      if (p + 2 > b.size()) return fail("truncated marker segment");
      const size_t len = size_t(u16(p)), seg = p + 2, end = p + len;
      if (len < 2 || end > b.size()) return fail("truncated marker segment");
      switch (m) {
        case 0xC0:
        case 0xC1:
          if (!readSof(seg, end)) return false;
          break;
        case 0xC2:
          progressive = true;
          if (!readSof(seg, end)) return false;
          break;
        case 0xC3: case 0xC5: case 0xC6: case 0xC7: case 0xC9: case 0xCA: case 0xCB: case 0xCD: case 0xCE:
        case 0xCF:
          return fail("lossless, hierarchical and arithmetic-coded JPEG are not supported");
        case 0xC4:
          if (!readDht(seg, end)) return false;
          break;
        case 0xDB:
          if (!readDqt(seg, end)) return false;
          break;
        case 0xDD:
          if (len < 4) return fail("bad restart interval");
          restartInterval = u16(seg);
          break;
        case 0xDA: {
          if (!frame) return fail("scan before the frame header");
          size_t after = 0;
          if (!scan(seg, end, after)) return false;
          p = after;
          continue;
        }
        case 0xE0:
          if (len >= 7 && std::memcmp(&b[seg], "JFIF", 4) == 0) jfif = true;
          break;
        case 0xEE:
          if (len >= 14 && std::memcmp(&b[seg], "Adobe", 5) == 0) {
            adobe = true;
            adobeTransform = b[seg + 11];
          }
          break;
        default:
          break;  // APPn, COM and others
      }
      p = end;
    }
    if (!frame) return fail("no frame");
    return true;
  }
};

// Triangle-filter 2x upsampling of one line (the 3:1 weights of the
// neighbouring samples), edges replicated.
void up2(const int* in, int n, int* out, int outN) {
  for (int i = 0; i < outN; ++i) {
    const int s = i >> 1;
    const int far = (i & 1) ? std::min(s + 1, n - 1) : std::max(s - 1, 0);
    out[i] = 3 * in[std::min(s, n - 1)] + in[far];
  }
}

}  // namespace

bool readJpeg(const std::vector<unsigned char>& bytes, float gamma, Image& im, std::string& err) {
  Decoder d(bytes, err);
  if (!d.run()) return false;
  const int W = d.width, H = d.height;
  // sample planes per component (padded block grid)
  std::vector<std::vector<uint8_t>> planes(d.comps.size());
  for (size_t ci = 0; ci < d.comps.size(); ++ci) {
    Component& c = d.comps[ci];
    const int pw = c.bw * 8;
    planes[ci].assign(size_t(pw) * size_t(c.bh) * 8, 0);
    for (int by = 0; by < c.bh; ++by)
      for (int bx = 0; bx < c.bw; ++bx)
        idct8x8(&c.coef[(size_t(by) * size_t(c.bw) + size_t(bx)) * 64], d.qt[c.tq],
                &planes[ci][(size_t(by) * 8) * size_t(pw) + size_t(bx) * 8], pw);
  }
  // component samples at full resolution
  std::vector<std::vector<int>> full(d.comps.size(), std::vector<int>(size_t(W) * size_t(H)));
  for (size_t ci = 0; ci < d.comps.size(); ++ci) {
    const Component& c = d.comps[ci];
    const int pw = c.bw * 8;
    const int cw = (W * c.h + d.hmax - 1) / d.hmax, ch = (H * c.v + d.vmax - 1) / d.vmax;
    const int sx = d.hmax / c.h, sy = d.vmax / c.v;
    const bool exact = d.hmax % c.h == 0 && d.vmax % c.v == 0;
    auto at = [&](int x, int y) { return int(planes[ci][size_t(y) * size_t(pw) + size_t(x)]); };
    std::vector<int>& o = full[ci];
    if (exact && sx == 1 && sy == 1) {
      for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) o[size_t(y) * W + x] = at(x, y);
    } else if (exact && sx == 2 && (sy == 1 || sy == 2)) {
      std::vector<int> row(static_cast<size_t>(cw)), line(static_cast<size_t>(2 * cw));
      for (int y = 0; y < H; ++y) {
        if (sy == 1) {
          for (int x = 0; x < cw; ++x) row[size_t(x)] = at(x, y);
          up2(row.data(), cw, line.data(), W);
          for (int x = 0; x < W; ++x) o[size_t(y) * W + x] = (line[size_t(x)] + 2) >> 2;
        } else {
          // vertical 3:1 blend of the nearer and the farther chroma row, then horizontal
          const int s = y >> 1, f = (y & 1) ? std::min(s + 1, ch - 1) : std::max(s - 1, 0);
          for (int x = 0; x < cw; ++x) row[size_t(x)] = 3 * at(x, std::min(s, ch - 1)) + at(x, f);
          up2(row.data(), cw, line.data(), W);
          for (int x = 0; x < W; ++x) o[size_t(y) * W + x] = (line[size_t(x)] + 8) >> 4;
        }
      }
    } else {
      for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
          o[size_t(y) * W + x] = at(std::min(x * c.h / d.hmax, cw - 1), std::min(y * c.v / d.vmax, ch - 1));
    }
  }
  // colour: YCbCr (JFIF) unless an Adobe marker says untransformed or the
  // component ids spell R, G, B
  const bool rgbIds = d.comps.size() == 3 && d.comps[0].id == 'R' && d.comps[1].id == 'G' && d.comps[2].id == 'B';
  const bool ycc = d.comps.size() == 3 && !(d.adobe && d.adobeTransform == 0) && !(rgbIds && !d.jfif && !d.adobe);
  im.w = uint32_t(W);
  im.h = uint32_t(H);
  im.rgba.assign(size_t(W) * size_t(H) * 4, 0.f);
  float lut[256];
  for (int v = 0; v < 256; ++v) lut[v] = float(std::pow(double(float(v) / 255.0f), double(gamma)));
  for (size_t i = 0; i < size_t(W) * size_t(H); ++i) {
    int r, g, bl;
    if (d.comps.size() == 1) {
      r = g = bl = full[0][i];
    } else if (ycc) {
      const double y = full[0][i], cb = full[1][i] - 128.0, cr = full[2][i] - 128.0;
      r = int(std::lround(y + 1.402 * cr));
      g = int(std::floor(y - 0.344136 * cb - 0.714136 * cr + 0.5));
      bl = int(std::lround(y + 1.772 * cb));
    } else {
      r = full[0][i], g = full[1][i], bl = full[2][i];
    }
    const int c8[3] = {std::min(255, std::max(0, r)), std::min(255, std::max(0, g)), std::min(255, std::max(0, bl))};
    for (int k = 0; k < 3; ++k) im.rgba[i * 4 + size_t(k)] = lut[c8[k]];
    im.rgba[i * 4 + 3] = 1.0f;
  }
  return true;
}

}  // namespace ash
