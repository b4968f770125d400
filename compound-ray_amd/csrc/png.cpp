#include "png.h"

#include <cstring>

namespace eye {

namespace {

// ---- DEFLATE (RFC 1951) ----------------------------------------------------

struct BitReader {
  const uint8_t* p;
  size_t n;
  size_t pos = 0;  // byte position
  uint32_t bitbuf = 0;
  int bitcnt = 0;
  bool overrun = false;

  BitReader(const uint8_t* data, size_t size) : p(data), n(size) {}

  uint32_t bits(int need) {
    while (bitcnt < need) {
      uint32_t byte = 0;
      if (pos < n)
        byte = p[pos++];
      else
        overrun = true;
      bitbuf |= byte << bitcnt;
      bitcnt += 8;
    }
    const uint32_t v = need == 0 ? 0u : (bitbuf & ((need == 32) ? 0xffffffffu : ((1u << need) - 1u)));
    bitbuf = need == 32 ? 0u : (bitbuf >> need);
    bitcnt -= need;
    return v;
  }
  void align_byte() {
    bitbuf = 0;
    bitcnt = 0;
  }
};

// Canonical Huffman code: number of codes of each length and the symbols in
// code order.  Decoding walks the code one bit at a time (codes are sent most
// significant bit first).
struct Huffman {
  uint16_t count[16];
  uint16_t symbol[320];
};

bool build_huffman(Huffman* h, const uint8_t* lengths, int n) {
  std::memset(h->count, 0, sizeof(h->count));
  for (int s = 0; s < n; ++s) h->count[lengths[s]]++;
  h->count[0] = 0;
  // over-subscribed sets are malformed; incomplete ones are allowed (a single
  // distance code is legal)
  int left = 1;
  for (int len = 1; len < 16; ++len) {
    left <<= 1;
    left -= h->count[len];
    if (left < 0) return false;
  }
  uint16_t offs[16];
  offs[1] = 0;
  for (int len = 1; len < 15; ++len) offs[len + 1] = (uint16_t)(offs[len] + h->count[len]);
  for (int s = 0; s < n; ++s)
    if (lengths[s] != 0) h->symbol[offs[lengths[s]]++] = (uint16_t)s;
  return true;
}

int decode_symbol(BitReader& br, const Huffman& h) {
  int code = 0, first = 0, index = 0;
  for (int len = 1; len < 16; ++len) {
    code |= (int)br.bits(1);
    const int count = h.count[len];
    if (code - count < first) return h.symbol[index + (code - first)];
    index += count;
    first += count;
    first <<= 1;
    code <<= 1;
  }
  return -1;
}

const uint16_t kLenBase[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                               31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t kDistBase[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,   33,   49,    65,    97,    129,
                                193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

bool inflate_codes(BitReader& br, const Huffman& lit, const Huffman& dist, std::vector<uint8_t>* out,
                   std::string* err) {
  for (;;) {
    const int sym = decode_symbol(br, lit);
    if (sym < 0 || br.overrun) {
      *err = "deflate: bad literal/length code";
      return false;
    }
    if (sym < 256) {
      out->push_back((uint8_t)sym);
      continue;
    }
    if (sym == 256) return true;
    const int li = sym - 257;
    if (li >= 29) {
      *err = "deflate: bad length symbol";
      return false;
    }
    const size_t len = kLenBase[li] + br.bits(kLenExtra[li]);
    const int ds = decode_symbol(br, dist);
    if (ds < 0 || ds >= 30) {
      *err = "deflate: bad distance code";
      return false;
    }
    const size_t d = kDistBase[ds] + br.bits(kDistExtra[ds]);
    if (d > out->size() || br.overrun) {
      *err = "deflate: distance beyond the output";
      return false;
    }
    const size_t from = out->size() - d;
    for (size_t k = 0; k < len; ++k) out->push_back((*out)[from + k]);
  }
}

bool inflate_raw(const uint8_t* data, size_t size, std::vector<uint8_t>* out, size_t* consumed, std::string* err) {
  BitReader br(data, size);
  Huffman fixed_lit, fixed_dist;
  {
    uint8_t l[288];
    for (int s = 0; s < 144; ++s) l[s] = 8;
    for (int s = 144; s < 256; ++s) l[s] = 9;
    for (int s = 256; s < 280; ++s) l[s] = 7;
    for (int s = 280; s < 288; ++s) l[s] = 8;
    build_huffman(&fixed_lit, l, 288);
    uint8_t d[30];
    for (int s = 0; s < 30; ++s) d[s] = 5;
    build_huffman(&fixed_dist, d, 30);
  }
  bool last = false;
  while (!last) {
    last = br.bits(1) != 0;
    const uint32_t type = br.bits(2);
    if (type == 0) {
      br.align_byte();
      if (br.pos + 4 > size) {
        *err = "deflate: truncated stored block";
        return false;
      }
      const uint32_t len = data[br.pos] | (uint32_t)data[br.pos + 1] << 8;
      const uint32_t nlen = data[br.pos + 2] | (uint32_t)data[br.pos + 3] << 8;
      br.pos += 4;
      if ((len ^ 0xffffu) != nlen || br.pos + len > size) {
        *err = "deflate: bad stored block";
        return false;
      }
      out->insert(out->end(), data + br.pos, data + br.pos + len);
      br.pos += len;
    } else if (type == 1) {
      if (!inflate_codes(br, fixed_lit, fixed_dist, out, err)) return false;
    } else if (type == 2) {
      const int nlit = (int)br.bits(5) + 257;
      const int ndist = (int)br.bits(5) + 1;
      const int ncode = (int)br.bits(4) + 4;
      if (nlit > 286 || ndist > 30) {
        *err = "deflate: bad code counts";
        return false;
      }
      static const uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
      uint8_t cl[19] = {0};
      for (int k = 0; k < ncode; ++k) cl[order[k]] = (uint8_t)br.bits(3);
      Huffman lencode;
      if (!build_huffman(&lencode, cl, 19)) {
        *err = "deflate: bad code-length code";
        return false;
      }
      uint8_t lengths[320] = {0};
      int idx = 0;
      while (idx < nlit + ndist) {
        const int sym = decode_symbol(br, lencode);
        if (sym < 0 || br.overrun) {
          *err = "deflate: bad code length";
          return false;
        }
        if (sym < 16) {
          lengths[idx++] = (uint8_t)sym;
          continue;
        }
        int rep = 0;
        uint8_t val = 0;
        if (sym == 16) {
          if (idx == 0) {
            *err = "deflate: repeat with no previous length";
            return false;
          }
          val = lengths[idx - 1];
          rep = 3 + (int)br.bits(2);
        } else if (sym == 17) {
          rep = 3 + (int)br.bits(3);
        } else {
          rep = 11 + (int)br.bits(7);
        }
        if (idx + rep > nlit + ndist) {
          *err = "deflate: code lengths overrun";
          return false;
        }
        while (rep-- > 0) lengths[idx++] = val;
      }
      if (lengths[256] == 0) {
        *err = "deflate: no end-of-block code";
        return false;
      }
      Huffman lit, dist;
      if (!build_huffman(&lit, lengths, nlit) || !build_huffman(&dist, lengths + nlit, ndist)) {
        *err = "deflate: bad literal/distance code";
        return false;
      }
      if (!inflate_codes(br, lit, dist, out, err)) return false;
    } else {
      *err = "deflate: reserved block type";
      return false;
    }
    if (br.overrun) {
      *err = "deflate: truncated stream";
      return false;
    }
  }
  *consumed = br.pos;
  return true;
}

uint32_t adler32(const uint8_t* p, size_t n) {
  uint32_t a = 1, b = 0;
  for (size_t i = 0; i < n; ++i) {
    a = (a + p[i]) % 65521u;
    b = (b + a) % 65521u;
  }
  return (b << 16) | a;
}

// ---- PNG ------------------------------------------------------------------

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

uint32_t crc32_of(const uint8_t* p, size_t n) {
  static uint32_t table[256];
  static bool ready = false;
  if (!ready) {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xedb88320u ^ (c >> 1) : c >> 1;
      table[i] = c;
    }
    ready = true;
  }
  uint32_t c = 0xffffffffu;
  for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xffu] ^ (c >> 8);
  return c ^ 0xffffffffu;
}

int paeth(int a, int b, int c) {
  const int p = a + b - c;
  const int pa = p > a ? p - a : a - p;
  const int pb = p > b ? p - b : b - p;
  const int pc = p > c ? p - c : c - p;
  if (pa <= pb && pa <= pc) return a;
  if (pb <= pc) return b;
  return c;
}

// Reverses the row filters of one (sub-)image in place; `rows` holds
// height x (1 + row_bytes) bytes.
bool unfilter(uint8_t* rows, int height, size_t row_bytes, int bpp, std::string* err) {
  uint8_t* prev = nullptr;
  for (int y = 0; y < height; ++y) {
    uint8_t* row = rows + (size_t)y * (row_bytes + 1);
    const int type = row[0];
    uint8_t* cur = row + 1;
    for (size_t i = 0; i < row_bytes; ++i) {
      const int a = i >= (size_t)bpp ? cur[i - bpp] : 0;
      const int b = prev ? prev[i] : 0;
      const int c = (prev && i >= (size_t)bpp) ? prev[i - bpp] : 0;
      int v = cur[i];
      switch (type) {
        case 0: break;
        case 1: v += a; break;
        case 2: v += b; break;
        case 3: v += (a + b) >> 1; break;
        case 4: v += paeth(a, b, c); break;
        default:
          *err = "png: unknown row filter " + std::to_string(type);
          return false;
      }
      cur[i] = (uint8_t)v;
    }
    prev = cur;
  }
  return true;
}

struct PngInfo {
  int w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
  int channels = 0;
  std::vector<uint8_t> palette;  // rgb triples
  std::vector<uint8_t> pal_alpha;
  bool has_trns = false;
  uint16_t trns[3] = {0, 0, 0};
};

uint32_t sample_at(const uint8_t* row, int depth, size_t index) {
  if (depth == 8) return row[index];
  if (depth == 16) return (uint32_t)row[2 * index] << 8 | row[2 * index + 1];
  const size_t bit = index * (size_t)depth;
  const int shift = 8 - depth - (int)(bit & 7u);
  return (row[bit >> 3] >> shift) & ((1u << depth) - 1u);
}

uint8_t to8(uint32_t v, int depth) {
  if (depth == 8) return (uint8_t)v;
  if (depth == 16) return (uint8_t)((v * 255u + 32767u) / 65535u);
  return (uint8_t)(v * 255u / ((1u << depth) - 1u));
}

// Writes pixel (x, y) of the full image from sample index `i` of `row`.
void store_pixel(const PngInfo& P, const uint8_t* row, size_t i, uint8_t* out) {
  const int d = P.depth;
  switch (P.ctype) {
    case 0: {
      const uint32_t g = sample_at(row, d, i);
      const uint8_t v = to8(g, d);
      out[0] = out[1] = out[2] = v;
      out[3] = (P.has_trns && g == P.trns[0]) ? 0 : 255;
      break;
    }
    case 2: {
      const uint32_t r = sample_at(row, d, 3 * i), g = sample_at(row, d, 3 * i + 1), b = sample_at(row, d, 3 * i + 2);
      out[0] = to8(r, d);
      out[1] = to8(g, d);
      out[2] = to8(b, d);
      out[3] = (P.has_trns && r == P.trns[0] && g == P.trns[1] && b == P.trns[2]) ? 0 : 255;
      break;
    }
    case 3: {
      const uint32_t k = sample_at(row, d, i);
      if (3 * (size_t)k + 2 < P.palette.size()) {
        out[0] = P.palette[3 * k];
        out[1] = P.palette[3 * k + 1];
        out[2] = P.palette[3 * k + 2];
      } else {
        out[0] = out[1] = out[2] = 0;
      }
      out[3] = k < P.pal_alpha.size() ? P.pal_alpha[k] : 255;
      break;
    }
    case 4: {
      const uint8_t v = to8(sample_at(row, d, 2 * i), d);
      out[0] = out[1] = out[2] = v;
      out[3] = to8(sample_at(row, d, 2 * i + 1), d);
      break;
    }
    default: {  // 6
      for (int c = 0; c < 4; ++c) out[c] = to8(sample_at(row, d, 4 * i + c), d);
      break;
    }
  }
}

}  // namespace

bool zlib_inflate(const uint8_t* data, size_t size, std::vector<uint8_t>* out, std::string* err) {
  if (size < 6) {
    *err = "zlib: stream too short";
    return false;
  }
  const uint32_t cmf = data[0], flg = data[1];
  if ((cmf & 15u) != 8u || (cmf >> 4) > 7u || ((cmf << 8) | flg) % 31u != 0u || (flg & 0x20u)) {
    *err = "zlib: bad header";
    return false;
  }
  size_t used = 0;
  out->clear();
  if (!inflate_raw(data + 2, size - 2, out, &used, err)) return false;
  if (2 + used + 4 > size) {
    *err = "zlib: missing checksum";
    return false;
  }
  if (be32(data + 2 + used) != adler32(out->data(), out->size())) {
    *err = "zlib: checksum mismatch";
    return false;
  }
  return true;
}

bool png_decode(const uint8_t* data, size_t size, int* width, int* height, std::vector<uint8_t>* rgba,
                std::string* err) {
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (size < 8 || std::memcmp(data, sig, 8) != 0) {
    *err = "png: not a PNG file";
    return false;
  }
  PngInfo P;
  std::vector<uint8_t> idat;
  bool seen_ihdr = false, seen_iend = false;
  size_t pos = 8;
  while (pos + 12 <= size && !seen_iend) {
    const uint32_t len = be32(data + pos);
    const uint8_t* type = data + pos + 4;
    if (len > size - pos - 12) {
      *err = "png: truncated chunk";
      return false;
    }
    const uint8_t* body = data + pos + 8;
    if (be32(body + len) != crc32_of(type, len + 4)) {
      *err = "png: chunk CRC mismatch";
      return false;
    }
    if (!std::memcmp(type, "IHDR", 4)) {
      if (len != 13) {
        *err = "png: bad IHDR";
        return false;
      }
      P.w = (int)be32(body);
      P.h = (int)be32(body + 4);
      P.depth = body[8];
      P.ctype = body[9];
      P.interlace = body[12];
      if (body[10] != 0 || body[11] != 0 || P.interlace > 1 || P.w <= 0 || P.h <= 0 || P.w > (1 << 15) ||
          P.h > (1 << 15)) {
        *err = "png: unsupported header";
        return false;
      }
      const int d = P.depth;
      const bool ok = (P.ctype == 0 && (d == 1 || d == 2 || d == 4 || d == 8 || d == 16)) ||
                      (P.ctype == 3 && (d == 1 || d == 2 || d == 4 || d == 8)) ||
                      ((P.ctype == 2 || P.ctype == 4 || P.ctype == 6) && (d == 8 || d == 16));
      if (!ok) {
        *err = "png: bad colour type / bit depth";
        return false;
      }
      P.channels = P.ctype == 0 ? 1 : P.ctype == 2 ? 3 : P.ctype == 3 ? 1 : P.ctype == 4 ? 2 : 4;
      seen_ihdr = true;
    } else if (!std::memcmp(type, "PLTE", 4)) {
      P.palette.assign(body, body + len);
    } else if (!std::memcmp(type, "tRNS", 4)) {
      if (P.ctype == 3) {
        P.pal_alpha.assign(body, body + len);
      } else if (P.ctype == 0 && len >= 2) {
        P.has_trns = true;
        P.trns[0] = (uint16_t)(body[0] << 8 | body[1]);
      } else if (P.ctype == 2 && len >= 6) {
        P.has_trns = true;
        for (int c = 0; c < 3; ++c) P.trns[c] = (uint16_t)(body[2 * c] << 8 | body[2 * c + 1]);
      }
    } else if (!std::memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), body, body + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      seen_iend = true;
    } else if (!(type[0] & 0x20)) {
      *err = std::string("png: unknown critical chunk ") + std::string((const char*)type, 4);
      return false;
    }
    pos += 12 + len;
  }
  if (!seen_ihdr || idat.empty()) {
    *err = "png: missing IHDR or IDAT";
    return false;
  }
  if (P.ctype == 3 && P.palette.empty()) {
    *err = "png: palette image without PLTE";
    return false;
  }
  std::vector<uint8_t> raw;
  if (!zlib_inflate(idat.data(), idat.size(), &raw, err)) return false;
  const int bits_pp = P.channels * P.depth;
  const int bpp = bits_pp >= 8 ? bits_pp / 8 : 1;
  rgba->assign((size_t)P.w * (size_t)P.h * 4u, 0);
  // passes: (x0, y0, dx, dy); one pass covering everything when not interlaced
  static const int adam7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                  {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
  static const int single[1][4] = {{0, 0, 1, 1}};
  const int npass = P.interlace ? 7 : 1;
  size_t off = 0;
  for (int pi = 0; pi < npass; ++pi) {
    const int* ps = P.interlace ? adam7[pi] : single[0];
    const int pw = (P.w - ps[0] + ps[2] - 1) / ps[2];
    const int ph = (P.h - ps[1] + ps[3] - 1) / ps[3];
    if (pw <= 0 || ph <= 0) continue;
    const size_t row_bytes = ((size_t)pw * (size_t)bits_pp + 7u) / 8u;
    const size_t need = (size_t)ph * (row_bytes + 1);
    if (off + need > raw.size()) {
      *err = "png: image data too short";
      return false;
    }
    uint8_t* rows = raw.data() + off;
    if (!unfilter(rows, ph, row_bytes, bpp, err)) return false;
    for (int y = 0; y < ph; ++y) {
      const uint8_t* row = rows + (size_t)y * (row_bytes + 1) + 1;
      const int fy = ps[1] + y * ps[3];
      for (int x = 0; x < pw; ++x) {
        const int fx = ps[0] + x * ps[2];
        store_pixel(P, row, (size_t)x, rgba->data() + ((size_t)fy * (size_t)P.w + (size_t)fx) * 4u);
      }
    }
    off += need;
  }
  *width = P.w;
  *height = P.h;
  return true;
}

}  // namespace eye
