#include "jpeg.h"

#include <cmath>
#include <cstring>

namespace eye {

namespace {

// Natural (row-major) index of the k-th coefficient in zig-zag order
// (ITU-T T.81, figure A.6).
constexpr uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Canonical Huffman table (T.81 annex C and F.2.2.3): the codes of each
// length are consecutive integers; mincode/maxcode bound them per length and
// valptr indexes the first of their values.
struct HuffTable {
  bool present = false;
  uint8_t vals[256];
  int32_t mincode[17], maxcode[17], valptr[17];
};

bool build_table(HuffTable* t, const uint8_t bits[16], const uint8_t* vals, int nvals) {
  int code = 0, k = 0;
  for (int l = 1; l <= 16; ++l) {
    const int c = bits[l - 1];
    t->valptr[l] = k;
    t->mincode[l] = code;
    code += c;
    k += c;
    t->maxcode[l] = c ? code - 1 : -1;
    if (code > (1 << l)) return false;   // over-subscribed code lengths
    code <<= 1;
  }
  if (k != nvals) return false;
  std::memcpy(t->vals, vals, (size_t)nvals);
  t->present = true;
  return true;
}

// Entropy-coded segment reader: removes the 0x00 stuffed after every 0xFF
// data byte and stops at a marker (it then feeds zero bits).
struct BitReader {
  const uint8_t* p;
  size_t n;
  size_t pos;
  uint32_t buf = 0;
  int cnt = 0;
  bool at_marker = false;
  bool exhausted = false;

  int bit() {
    if (cnt == 0) {
      uint32_t byte = 0;
      if (!at_marker && pos < n) {
        byte = p[pos];
        if (byte == 0xFF) {
          const uint32_t next = pos + 1 < n ? p[pos + 1] : 0xD9u;
          if (next == 0x00) {
            pos += 2;
          } else {
            at_marker = true;
            byte = 0;
          }
        } else {
          ++pos;
        }
      } else {
        exhausted = true;
      }
      buf = byte;
      cnt = 8;
    }
    --cnt;
    return (int)((buf >> cnt) & 1u);
  }
  int receive(int s) {
    int v = 0;
    for (int i = 0; i < s; ++i) v = (v << 1) | bit();
    return v;
  }
  // F.2.2.1 EXTEND: the s-bit magnitude category value as a signed integer
  static int extend(int v, int s) { return s == 0 ? 0 : (v < (1 << (s - 1)) ? v - (1 << s) + 1 : v); }
  bool decode(const HuffTable& t, int* out) {
    int code = 0;
    for (int l = 1; l <= 16; ++l) {
      code = (code << 1) | bit();
      if (t.maxcode[l] >= 0 && code <= t.maxcode[l] && code >= t.mincode[l]) {
        const int idx = t.valptr[l] + code - t.mincode[l];
        if (idx < 0 || idx > 255) return false;
        *out = t.vals[idx];
        return true;
      }
    }
    return false;
  }
  // Restart interval boundary: drop the remaining bits and step over RSTn.
  bool restart() {
    buf = 0;
    cnt = 0;
    while (pos < n && p[pos] != 0xFF) ++pos;   // tolerate padding before the marker
    if (pos + 1 < n && p[pos] == 0xFF && p[pos + 1] >= 0xD0 && p[pos + 1] <= 0xD7) {
      pos += 2;
      at_marker = false;
      return true;
    }
    return false;
  }
};

struct Component {
  int id = 0, h = 1, v = 1, tq = 0;
  int td = 0, ta = 0;
  int bw = 0, bh = 0;          // blocks per line and per column, padded to whole MCUs
  int dc = 0;                  // DC prediction
  bool decoded = false;
  std::vector<int16_t> coef;   // bw * bh blocks of 64 quantised coefficients (natural order)
};

struct Decoder {
  const uint8_t* p;
  size_t n;
  std::string* err;
  uint16_t qt[4][64];
  bool qt_present[4] = {false, false, false, false};
  HuffTable dc[4], ac[4];
  Component comp[4];
  int ncomp = 0, width = 0, height = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
  int restart_interval = 0;
  int adobe_transform = -1;
  bool frame = false;

  bool fail(const std::string& m) {
    if (err) *err = "jpeg: " + m;
    return false;
  }

  bool read_dqt(const uint8_t* s, size_t len) {
    size_t i = 0;
    while (i < len) {
      const int pq = s[i] >> 4, tq = s[i] & 15;
      ++i;
      if (tq > 3 || pq > 1) return fail("bad quantisation table");
      const size_t need = pq ? 128 : 64;
      if (i + need > len) return fail("truncated quantisation table");
      for (int k = 0; k < 64; ++k)
        qt[tq][kZigzag[k]] = pq ? (uint16_t)(s[i + 2 * k] << 8 | s[i + 2 * k + 1]) : s[i + k];
      qt_present[tq] = true;
      i += need;
    }
    return true;
  }

  bool read_dht(const uint8_t* s, size_t len) {
    size_t i = 0;
    while (i < len) {
      if (i + 17 > len) return fail("truncated Huffman table");
      const int tc = s[i] >> 4, th = s[i] & 15;
      if (tc > 1 || th > 3) return fail("bad Huffman table class or slot");
      uint8_t bits[16];
      int total = 0;
      for (int l = 0; l < 16; ++l) {
        bits[l] = s[i + 1 + l];
        total += bits[l];
      }
      i += 17;
      if (total > 256 || i + (size_t)total > len) return fail("truncated Huffman table");
      if (!build_table(tc ? &ac[th] : &dc[th], bits, s + i, total)) return fail("invalid Huffman code lengths");
      i += (size_t)total;
    }
    return true;
  }

  bool read_sof(const uint8_t* s, size_t len) {
    if (frame) return fail("second frame header");
    if (len < 6) return fail("truncated frame header");
    if (s[0] != 8) return fail("only 8-bit samples are supported");
    height = s[1] << 8 | s[2];
    width = s[3] << 8 | s[4];
    ncomp = s[5];
    if (width == 0 || height == 0) return fail("zero or deferred (DNL) image size");
    if ((uint64_t)width * (uint64_t)height > (1ull << 28)) return fail("image too large");
    if (ncomp != 1 && ncomp != 3) return fail("only 1- and 3-component images are supported");
    if (len < 6 + 3 * (size_t)ncomp) return fail("truncated frame header");
    for (int c = 0; c < ncomp; ++c) {
      Component& k = comp[c];
      k.id = s[6 + 3 * c];
      k.h = s[7 + 3 * c] >> 4;
      k.v = s[7 + 3 * c] & 15;
      k.tq = s[8 + 3 * c];
      if (k.h < 1 || k.h > 4 || k.v < 1 || k.v > 4 || k.tq > 3) return fail("bad component parameters");
      hmax = k.h > hmax ? k.h : hmax;
      vmax = k.v > vmax ? k.v : vmax;
    }
    for (int c = 0; c < ncomp; ++c)
      if (hmax % comp[c].h || vmax % comp[c].v) return fail("non-integer chroma subsampling is not supported");
    mcux = (width + 8 * hmax - 1) / (8 * hmax);
    mcuy = (height + 8 * vmax - 1) / (8 * vmax);
    for (int c = 0; c < ncomp; ++c) {
      Component& k = comp[c];
      k.bw = mcux * k.h;
      k.bh = mcuy * k.v;
      k.coef.assign((size_t)k.bw * (size_t)k.bh * 64, 0);
    }
    frame = true;
    return true;
  }

  bool decode_block(BitReader& br, Component& k, int bx, int by) {
    int16_t* blk = &k.coef[((size_t)by * (size_t)k.bw + (size_t)bx) * 64];
    int t = 0;
    if (!br.decode(dc[k.td], &t) || t > 11) return fail("bad DC code");
    k.dc += BitReader::extend(br.receive(t), t);
    blk[0] = (int16_t)k.dc;
    for (int i = 1; i < 64;) {
      int rs = 0;
      if (!br.decode(ac[k.ta], &rs)) return fail("bad AC code");
      const int r = rs >> 4, sz = rs & 15;
      if (sz == 0) {
        if (r != 15) break;   // end of block
        i += 16;
        continue;
      }
      i += r;
      if (i > 63) return fail("AC coefficients run past the block");
      blk[kZigzag[i]] = (int16_t)BitReader::extend(br.receive(sz), sz);
      ++i;
    }
    return true;
  }

  // Decodes one scan whose entropy-coded data starts at *pos; leaves *pos at
  // the marker after it.
  bool read_sos(const uint8_t* s, size_t len, size_t* pos) {
    if (!frame) return fail("scan before the frame header");
    if (len < 1) return fail("truncated scan header");
    const int ns = s[0];
    if (ns < 1 || ns > ncomp || len < 4 + 2 * (size_t)ns) return fail("bad scan header");
    Component* sc[4];
    for (int j = 0; j < ns; ++j) {
      const int id = s[1 + 2 * j];
      Component* k = nullptr;
      for (int c = 0; c < ncomp; ++c)
        if (comp[c].id == id) k = &comp[c];
      if (!k) return fail("scan names an unknown component");
      for (int q = 0; q < j; ++q)
        if (sc[q] == k) return fail("scan names a component twice");
      k->td = s[2 + 2 * j] >> 4;
      k->ta = s[2 + 2 * j] & 15;
      if (k->td > 3 || k->ta > 3 || !dc[k->td].present || !ac[k->ta].present) return fail("missing Huffman table");
      if (!qt_present[k->tq]) return fail("missing quantisation table");
      if (k->decoded) return fail("component coded in two scans");
      sc[j] = k;
    }
    const uint8_t ss = s[1 + 2 * ns], se = s[2 + 2 * ns];
    if (ss != 0 || se != 63) return fail("spectral selection of a progressive scan");
    BitReader br{p, n, *pos};
    for (int j = 0; j < ns; ++j) sc[j]->dc = 0;
    int todo = restart_interval;
    auto next_mcu = [&]() -> bool {
      if (restart_interval == 0) return true;
      if (--todo > 0) return true;
      todo = restart_interval;
      if (!br.restart()) return fail("missing restart marker");
      for (int j = 0; j < ns; ++j) sc[j]->dc = 0;
      return true;
    };
    if (ns == 1) {
      // non-interleaved: the component's own blocks in raster order
      Component& k = *sc[0];
      const int cw = (width * k.h + hmax - 1) / hmax, ch = (height * k.v + vmax - 1) / vmax;
      const int nbx = (cw + 7) / 8, nby = (ch + 7) / 8;
      for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx) {
          if (!decode_block(br, k, bx, by)) return false;
          if ((bx + 1 < nbx || by + 1 < nby) && !next_mcu()) return false;
        }
    } else {
      for (int my = 0; my < mcuy; ++my)
        for (int mx = 0; mx < mcux; ++mx) {
          for (int j = 0; j < ns; ++j) {
            Component& k = *sc[j];
            for (int v = 0; v < k.v; ++v)
              for (int h = 0; h < k.h; ++h)
                if (!decode_block(br, k, mx * k.h + h, my * k.v + v)) return false;
          }
          if ((mx + 1 < mcux || my + 1 < mcuy) && !next_mcu()) return false;
        }
    }
    if (br.exhausted) return fail("entropy-coded data ends early");
    for (int j = 0; j < ns; ++j) sc[j]->decoded = true;
    // step to the next marker
    size_t q = br.pos;
    while (q + 1 < n && !(p[q] == 0xFF && p[q + 1] != 0x00 && !(p[q + 1] >= 0xD0 && p[q + 1] <= 0xD7))) ++q;
    *pos = q;
    return true;
  }

  bool parse() {
    if (!is_jpeg(p, n)) return fail("not a JPEG file");
    size_t pos = 2;
    while (true) {
      if (pos + 1 >= n) return fail("missing end-of-image marker");
      if (p[pos] != 0xFF) return fail("marker expected");
      while (pos + 1 < n && p[pos + 1] == 0xFF) ++pos;   // fill bytes
      if (pos + 1 >= n) return fail("truncated marker");
      const uint8_t m = p[pos + 1];
      pos += 2;
      if (m == 0xD9) break;
      if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) continue;
      if (m == 0xD8) return fail("nested start-of-image");
      if (pos + 2 > n) return fail("truncated segment");
      const size_t len = (size_t)(p[pos] << 8 | p[pos + 1]);
      if (len < 2 || pos + len > n) return fail("segment runs past the end of the file");
      const uint8_t* s = p + pos + 2;
      const size_t sl = len - 2;
      pos += len;
      switch (m) {
        case 0xDB:
          if (!read_dqt(s, sl)) return false;
          break;
        case 0xC4:
          if (!read_dht(s, sl)) return false;
          break;
        case 0xC0:
        case 0xC1:
          if (!read_sof(s, sl)) return false;
          break;
        case 0xC2:
        case 0xC6:
        case 0xCA:
        case 0xCE:
          return fail("progressive JPEG is not supported");
        case 0xC3:
        case 0xC5:
        case 0xC7:
        case 0xC9:
        case 0xCB:
        case 0xCD:
        case 0xCF:
          return fail("lossless, hierarchical or arithmetic-coded JPEG is not supported");
        case 0xCC:
          return fail("arithmetic coding is not supported");
        case 0xDD:
          if (sl < 2) return fail("truncated restart interval");
          restart_interval = s[0] << 8 | s[1];
          break;
        case 0xEE:
          if (sl >= 12 && std::memcmp(s, "Adobe", 5) == 0) adobe_transform = s[11];
          break;
        case 0xDA:
          if (!read_sos(s, sl, &pos)) return false;
          break;
        default:
          break;   // APPn, COM and others carry nothing the image needs
      }
    }
    if (!frame) return fail("no frame header");
    for (int c = 0; c < ncomp; ++c)
      if (!comp[c].decoded) return fail("a component has no scan");
    return true;
  }
};

// c[x][u] = C(u) / 2 * cos((2x + 1) u pi / 16), C(0) = 1 / sqrt(2), C(u) = 1.
struct IdctTable {
  float c[8][8];
  IdctTable() {
    const double pi = 3.14159265358979323846;
    for (int x = 0; x < 8; ++x)
      for (int u = 0; u < 8; ++u)
        c[x][u] = (float)((u == 0 ? std::sqrt(0.5) : 1.0) * 0.5 * std::cos((2 * x + 1) * u * pi / 16.0));
  }
};

// Dequantised inverse DCT of one block into 8x8 samples (level shift,
// rounding to nearest, clamping to [0, 255]).
void idct_block(const int16_t* coef, const uint16_t* q, uint8_t* out, int stride) {
  static const IdctTable T;
  float f[64], tmp[64];
  for (int i = 0; i < 64; ++i) f[i] = (float)coef[i] * (float)q[i];
  for (int v = 0; v < 8; ++v)
    for (int x = 0; x < 8; ++x) {
      float s = 0.0f;
      for (int u = 0; u < 8; ++u) s += T.c[x][u] * f[v * 8 + u];
      tmp[v * 8 + x] = s;
    }
  for (int y = 0; y < 8; ++y)
    for (int x = 0; x < 8; ++x) {
      float s = 0.0f;
      for (int v = 0; v < 8; ++v) s += T.c[y][v] * tmp[v * 8 + x];
      const float val = std::floor(s + 128.0f + 0.5f);
      out[y * stride + x] = (uint8_t)(val < 0.0f ? 0.0f : (val > 255.0f ? 255.0f : val));
    }
}

uint8_t clamp_round(float v) {
  const float r = std::floor(v + 0.5f);
  return (uint8_t)(r < 0.0f ? 0.0f : (r > 255.0f ? 255.0f : r));
}

}  // namespace

bool is_jpeg(const uint8_t* data, size_t size) {
  return size >= 3 && data[0] == 0xFF && data[1] == 0xD8 && data[2] == 0xFF;
}

bool jpeg_decode(const uint8_t* data, size_t size, int* width, int* height, std::vector<uint8_t>* rgba,
                 std::string* err) {
  Decoder d;
  d.p = data;
  d.n = size;
  d.err = err;
  if (!d.parse()) return false;
  const int W = d.width, H = d.height;
  // component planes at their own resolution (padded to whole MCUs)
  std::vector<std::vector<uint8_t>> plane((size_t)d.ncomp);
  for (int c = 0; c < d.ncomp; ++c) {
    const Component& k = d.comp[c];
    const size_t pw = (size_t)k.bw * 8, ph = (size_t)k.bh * 8;
    plane[(size_t)c].assign(pw * ph, 0);
    for (int by = 0; by < k.bh; ++by)
      for (int bx = 0; bx < k.bw; ++bx)
        idct_block(&k.coef[((size_t)by * (size_t)k.bw + (size_t)bx) * 64], d.qt[k.tq],
                   &plane[(size_t)c][(size_t)by * 8 * pw + (size_t)bx * 8], (int)pw);
  }
  // every component at full resolution, as 8-bit samples.  Factors
  // (fx, fy) = (hmax / h, vmax / v) of (2, 1), (1, 2) and (2, 2) use the
  // triangle filter: sample centres sit at (j + 0.5) * f - 0.5 and pixels are
  // interpolated linearly between the two nearest along each halved axis,
  // clamped at the component's true edge.  Other factors replicate each
  // sample over its f x f pixels.
  std::vector<uint8_t> full((size_t)d.ncomp * (size_t)W * (size_t)H);
  for (int c = 0; c < d.ncomp; ++c) {
    const Component& k = d.comp[c];
    const int fx = d.hmax / k.h, fy = d.vmax / k.v;
    const int cw = (W * k.h + d.hmax - 1) / d.hmax, ch = (H * k.v + d.vmax - 1) / d.vmax;
    const bool triangle = fx <= 2 && fy <= 2;
    const size_t pw = (size_t)k.bw * 8;
    const uint8_t* src = plane[(size_t)c].data();
    uint8_t* dst = &full[(size_t)c * (size_t)W * (size_t)H];
    for (int y = 0; y < H; ++y) {
      int y0 = y / fy, y1 = y0;
      float wy = 0.0f;
      if (triangle && fy == 2) {
        const float ys = (y + 0.5f) * 0.5f - 0.5f;
        y0 = (int)std::floor(ys);
        wy = ys - (float)y0;
        y1 = y0 + 1;
      }
      y0 = y0 < 0 ? 0 : (y0 >= ch ? ch - 1 : y0);
      y1 = y1 < 0 ? 0 : (y1 >= ch ? ch - 1 : y1);
      for (int x = 0; x < W; ++x) {
        int x0 = x / fx, x1 = x0;
        float wx = 0.0f;
        if (triangle && fx == 2) {
          const float xs = (x + 0.5f) * 0.5f - 0.5f;
          x0 = (int)std::floor(xs);
          wx = xs - (float)x0;
          x1 = x0 + 1;
        }
        x0 = x0 < 0 ? 0 : (x0 >= cw ? cw - 1 : x0);
        x1 = x1 < 0 ? 0 : (x1 >= cw ? cw - 1 : x1);
        const float a = src[(size_t)y0 * pw + (size_t)x0], b = src[(size_t)y0 * pw + (size_t)x1];
        const float e = src[(size_t)y1 * pw + (size_t)x0], f = src[(size_t)y1 * pw + (size_t)x1];
        dst[(size_t)y * W + x] = clamp_round((1.0f - wy) * ((1.0f - wx) * a + wx * b) + wy * ((1.0f - wx) * e + wx * f));
      }
    }
  }
  rgba->assign((size_t)W * (size_t)H * 4, 255);
  const size_t np = (size_t)W * (size_t)H;
  const bool ycc = d.ncomp == 3 && d.adobe_transform != 0;
  for (size_t i = 0; i < np; ++i) {
    uint8_t* o = &(*rgba)[4 * i];
    if (d.ncomp == 1) {
      o[0] = o[1] = o[2] = full[i];
    } else if (!ycc) {
      o[0] = full[i];
      o[1] = full[np + i];
      o[2] = full[2 * np + i];
    } else {
      // JFIF YCbCr -> RGB
      const float Y = full[i], cb = (float)full[np + i] - 128.0f, cr = (float)full[2 * np + i] - 128.0f;
      o[0] = clamp_round(Y + 1.402f * cr);
      o[1] = clamp_round(Y - 0.344136f * cb - 0.714136f * cr);
      o[2] = clamp_round(Y + 1.772f * cb);
    }
  }
  *width = W;
  *height = H;
  return true;
}

}  // namespace eye
