#include "scene_loader.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

#include "jpeg.h"
#include "json.h"
#include "png.h"

namespace eye {
namespace {

// Row-major 4x4 float matrix.
struct Mat4 {
  float m[16];
  static Mat4 identity() {
    Mat4 r;
    for (int i = 0; i < 16; ++i) r.m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    return r;
  }
  bool is_identity() const {
    for (int i = 0; i < 16; ++i)
      if (m[i] != ((i % 5 == 0) ? 1.0f : 0.0f)) return false;
    return true;
  }
};

Mat4 mul(const Mat4& a, const Mat4& b) {
  Mat4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      float s = 0.0f;
      for (int k = 0; k < 4; ++k) s = s + a.m[4 * i + k] * b.m[4 * k + j];
      r.m[4 * i + j] = s;
    }
  return r;
}

void apply(const Mat4& a, const float v[4], float out[4]) {
  for (int i = 0; i < 4; ++i)
    out[i] = ((a.m[4 * i] * v[0] + a.m[4 * i + 1] * v[1]) + a.m[4 * i + 2] * v[2]) + a.m[4 * i + 3] * v[3];
}

// Rotation matrix of a unit quaternion (w, x, y, z).
Mat4 quat_matrix(float w, float x, float y, float z) {
  Mat4 r = Mat4::identity();
  r.m[0] = 1.0f - 2.0f * (y * y + z * z);
  r.m[1] = 2.0f * (x * y - w * z);
  r.m[2] = 2.0f * (x * z + w * y);
  r.m[4] = 2.0f * (x * y + w * z);
  r.m[5] = 1.0f - 2.0f * (x * x + z * z);
  r.m[6] = 2.0f * (y * z - w * x);
  r.m[8] = 2.0f * (x * z - w * y);
  r.m[9] = 2.0f * (y * z + w * x);
  r.m[10] = 1.0f - 2.0f * (x * x + y * y);
  return r;
}

bool read_file(const std::string& path, std::string* out) {
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  *out = ss.str();
  return true;
}

int b64_value(char c) {
  if (c >= 'A' && c <= 'Z') return c - 'A';
  if (c >= 'a' && c <= 'z') return c - 'a' + 26;
  if (c >= '0' && c <= '9') return c - '0' + 52;
  if (c == '+' || c == '-') return 62;
  if (c == '/' || c == '_') return 63;
  return -1;
}

bool b64_decode(const char* s, size_t n, std::vector<uint8_t>* out) {
  out->clear();
  out->reserve(n / 4 * 3);
  uint32_t acc = 0;
  int bits = 0;
  for (size_t i = 0; i < n; ++i) {
    const char c = s[i];
    if (c == '=') break;
    if (c == '\n' || c == '\r' || c == ' ') continue;
    const int v = b64_value(c);
    if (v < 0) return false;
    acc = (acc << 6) | (uint32_t)v;
    bits += 6;
    if (bits >= 8) {
      bits -= 8;
      out->push_back((uint8_t)((acc >> bits) & 0xffu));
    }
  }
  return true;
}

bool extras_true(const JsonValue& extras, const char* key) {
  const JsonValue& v = extras[key];
  if (v.is_bool()) return v.b;
  if (v.is_string()) {
    std::string s = v.str;
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s == "true";
  }
  return false;
}

std::vector<std::string> split_on_space(const std::string& s) {
  std::vector<std::string> out;
  size_t last = 0;
  size_t pos = s.find(' ', 0);
  while (pos != std::string::npos) {
    if (pos != last) out.push_back(s.substr(last, pos - last));
    last = pos + 1;
    pos = s.find(' ', last);
  }
  out.push_back(s.substr(last));
  return out;
}

int components_of(const std::string& type) {
  if (type == "SCALAR") return 1;
  if (type == "VEC2") return 2;
  if (type == "VEC3") return 3;
  if (type == "VEC4") return 4;
  if (type == "MAT2") return 4;
  if (type == "MAT3") return 9;
  if (type == "MAT4") return 16;
  return 0;
}

int component_bytes(int ct) {
  switch (ct) {
    case 5120: case 5121: return 1;
    case 5122: case 5123: return 2;
    case 5125: case 5126: return 4;
    default: return 0;
  }
}

struct Loader {
  JsonValue doc;
  std::string dir;
  std::vector<std::vector<uint8_t>> buffers;
  HostScene* out = nullptr;
  std::string err;
  std::vector<int> texture_slot;   // glTF texture -> HostScene::textures index (-1: unusable, -2: not yet seen)
  std::vector<int> image_first;    // glTF image -> first texel (-1: not decoded yet)
  std::vector<int> image_w, image_h;

  struct Accessor {
    const uint8_t* base = nullptr;
    size_t count = 0;
    size_t stride = 0;
    int ncomp = 0;
    int ctype = 0;
    bool normalized = false;
  };

  bool load_buffers() {
    const JsonValue& bufs = doc["buffers"];
    for (size_t i = 0; i < bufs.size(); ++i) {
      const JsonValue& b = bufs[i];
      const std::string uri = b["uri"].as_string();
      std::vector<uint8_t> data;
      if (uri.rfind("data:", 0) == 0) {
        const size_t comma = uri.find(',');
        if (comma == std::string::npos || uri.find(";base64") == std::string::npos ||
            !b64_decode(uri.data() + comma + 1, uri.size() - comma - 1, &data)) {
          err = "buffer " + std::to_string(i) + ": unsupported data URI";
          return false;
        }
      } else if (!uri.empty()) {
        std::string bytes;
        if (!read_file(dir + uri, &bytes) && !read_file(uri, &bytes)) {
          err = "buffer " + std::to_string(i) + ": cannot open '" + uri + "'";
          return false;
        }
        data.assign(bytes.begin(), bytes.end());
      } else {
        err = "buffer " + std::to_string(i) + ": no uri (GLB files are not supported)";
        return false;
      }
      const size_t declared = (size_t)b["byteLength"].as_number(0.0);
      if (data.size() < declared) {
        err = "buffer " + std::to_string(i) + ": shorter than its byteLength";
        return false;
      }
      buffers.push_back(std::move(data));
    }
    return true;
  }

  // A byte count or offset from the JSON: absent gives 0; negative,
  // fractional, non-numeric or huge (2^48 and above) values are rejected, so
  // the bounds arithmetic below cannot wrap.
  static bool json_size(const JsonValue& v, size_t* out) {
    if (v.is_null()) {
      *out = 0;
      return true;
    }
    if (!v.is_number()) return false;
    const double d = v.num;
    if (!(d >= 0.0) || d != std::floor(d) || d >= 281474976710656.0) return false;
    *out = (size_t)d;
    return true;
  }

  bool accessor(int idx, Accessor* a) {
    const JsonValue& acc = doc["accessors"][(size_t)idx];
    if (!acc.is_object()) {
      err = "accessor " + std::to_string(idx) + " missing";
      return false;
    }
    if (!json_size(acc["count"], &a->count)) {
      err = "accessor " + std::to_string(idx) + ": count must be a non-negative integer";
      return false;
    }
    a->ncomp = components_of(acc["type"].as_string());
    a->ctype = acc["componentType"].as_int(0);
    a->normalized = acc["normalized"].is_bool() && acc["normalized"].b;
    const int cb = component_bytes(a->ctype);
    if (a->ncomp == 0 || cb == 0) {
      err = "accessor " + std::to_string(idx) + ": component type not supported";
      return false;
    }
    const int bv_idx = acc["bufferView"].as_int(-1);
    if (bv_idx < 0) {
      err = "accessor " + std::to_string(idx) + ": sparse/bufferless accessors are not supported";
      return false;
    }
    const JsonValue& bv = doc["bufferViews"][(size_t)bv_idx];
    const int buf = bv["buffer"].as_int(-1);
    if (buf < 0 || (size_t)buf >= buffers.size()) {
      err = "bufferView " + std::to_string(bv_idx) + ": bad buffer";
      return false;
    }
    size_t bv_off = 0, bv_len = 0, off = 0;
    if (!json_size(bv["byteOffset"], &bv_off) || !json_size(acc["byteOffset"], &off) ||
        !json_size(bv["byteStride"], &a->stride) || !json_size(bv["byteLength"], &bv_len)) {
      err = "accessor " + std::to_string(idx) + ": byte offsets, lengths and stride must be non-negative integers";
      return false;
    }
    const size_t size = buffers[(size_t)buf].size();
    if (bv_off > size || (!bv["byteLength"].is_null() && bv_len > size - bv_off)) {
      err = "bufferView " + std::to_string(bv_idx) + ": reads past the end of its buffer";
      return false;
    }
    if (bv["byteLength"].is_null()) bv_len = size - bv_off;
    const size_t elem = (size_t)cb * (size_t)a->ncomp;
    if (a->stride == 0) a->stride = elem;
    // off + stride * (count - 1) + elem <= the view's length, without overflow
    if (a->count > 0 && (off > bv_len || elem > bv_len - off || a->count - 1 > (bv_len - off - elem) / a->stride)) {
      err = "accessor " + std::to_string(idx) + ": reads past the end of its buffer view";
      return false;
    }
    a->base = buffers[(size_t)buf].data() + bv_off + off;
    return true;
  }

  static uint32_t get_uint(const Accessor& a, size_t i, int c) {
    const uint8_t* p = a.base + a.stride * i + (size_t)component_bytes(a.ctype) * c;
    switch (a.ctype) {
      case 5121: return *p;
      case 5123: { uint16_t v; memcpy(&v, p, 2); return v; }
      case 5125: { uint32_t v; memcpy(&v, p, 4); return v; }
      case 5120: return (uint32_t)(int32_t)(int8_t)*p;
      case 5122: { int16_t v; memcpy(&v, p, 2); return (uint32_t)(int32_t)v; }
      default: { float f; memcpy(&f, p, 4); return (uint32_t)f; }
    }
  }

  static float get_float(const Accessor& a, size_t i, int c) {
    const uint8_t* p = a.base + a.stride * i + (size_t)component_bytes(a.ctype) * c;
    float f;
    memcpy(&f, p, 4);
    return f;
  }

  // Vertex colour decode of cuda/LocalGeometry.h:106-146 (u8 / 255, u16 / 65535, f32 as is).
  static float get_colour(const Accessor& a, size_t i, int c) {
    const uint8_t* p = a.base + a.stride * i + (size_t)component_bytes(a.ctype) * c;
    switch (a.ctype) {
      case 5121: return (float)(*p) / 255.0f;
      case 5123: { uint16_t v; memcpy(&v, p, 2); return (float)v / 65535.0f; }
      default: { float f; memcpy(&f, p, 4); return f; }
    }
  }

  // Bytes of glTF image `idx`: a bufferView, a base64 data URI or a file
  // next to the glTF.
  bool image_bytes(int idx, std::vector<uint8_t>* bytes) {
    const JsonValue& im = doc["images"][(size_t)idx];
    const int bv_idx = im["bufferView"].as_int(-1);
    if (bv_idx >= 0) {
      const JsonValue& bv = doc["bufferViews"][(size_t)bv_idx];
      const int buf = bv["buffer"].as_int(-1);
      size_t off = 0, len = 0;
      if (buf < 0 || (size_t)buf >= buffers.size() || !json_size(bv["byteOffset"], &off) ||
          !json_size(bv["byteLength"], &len) || off > buffers[(size_t)buf].size() ||
          len > buffers[(size_t)buf].size() - off) {
        err = "image " + std::to_string(idx) + ": bad bufferView";
        return false;
      }
      bytes->assign(buffers[(size_t)buf].begin() + (ptrdiff_t)off, buffers[(size_t)buf].begin() + (ptrdiff_t)(off + len));
      return true;
    }
    const std::string uri = im["uri"].as_string();
    if (uri.rfind("data:", 0) == 0) {
      const size_t comma = uri.find(',');
      if (comma == std::string::npos || !b64_decode(uri.data() + comma + 1, uri.size() - comma - 1, bytes)) {
        err = "image " + std::to_string(idx) + ": unsupported data URI";
        return false;
      }
      return true;
    }
    std::string file;
    if (uri.empty() || (!read_file(dir + uri, &file) && !read_file(uri, &file))) {
      err = "image " + std::to_string(idx) + ": cannot open '" + uri + "'";
      return false;
    }
    bytes->assign(file.begin(), file.end());
    return true;
  }

  // HostScene texture index for glTF texture `t`, decoding its image on first
  // use; -1 (with a warning) when it cannot be sampled.  A texture without a
  // sampler repeats and filters linearly (MulticamScene.cpp:606-610); with
  // one, minFilter NEAREST selects point sampling (:619-621).
  int texture(int t) {
    const JsonValue& texs = doc["textures"];
    if (t < 0 || (size_t)t >= texs.size()) return -1;
    if (texture_slot.empty()) texture_slot.assign(texs.size(), -2);
    if (texture_slot[(size_t)t] != -2) return texture_slot[(size_t)t];
    texture_slot[(size_t)t] = -1;
    const JsonValue& tx = texs[(size_t)t];
    const int src = tx["source"].as_int(-1);
    const JsonValue& imgs = doc["images"];
    if (src < 0 || (size_t)src >= imgs.size()) {
      out->warnings.push_back("texture " + std::to_string(t) + ": no image");
      return -1;
    }
    if (image_first.empty()) {
      image_first.assign(imgs.size(), -1);
      image_w.assign(imgs.size(), 0);
      image_h.assign(imgs.size(), 0);
    }
    if (image_first[(size_t)src] < 0) {
      std::vector<uint8_t> bytes, rgba;
      std::string perr;
      int w = 0, h = 0;
      if (!image_bytes(src, &bytes)) {
        out->warnings.push_back(err + "; texture " + std::to_string(t) + " not sampled");
        err.clear();
        return -1;
      }
      const bool jpeg = is_jpeg(bytes.data(), bytes.size());
      const bool ok = jpeg ? jpeg_decode(bytes.data(), bytes.size(), &w, &h, &rgba, &perr)
                           : png_decode(bytes.data(), bytes.size(), &w, &h, &rgba, &perr);
      if (!ok) {
        out->warnings.push_back("image " + std::to_string(src) + ": " + perr +
                                (jpeg ? "" : " (PNG and JPEG images are decoded)") + "; texture " +
                                std::to_string(t) + " not sampled");
        return -1;
      }
      if (out->texels.size() + (size_t)w * (size_t)h > 0x7fffffffu) {
        out->warnings.push_back("image " + std::to_string(src) + ": texel budget exceeded, not sampled");
        return -1;
      }
      image_first[(size_t)src] = (int)out->texels.size();
      image_w[(size_t)src] = w;
      image_h[(size_t)src] = h;
      for (size_t k = 0; k < (size_t)w * (size_t)h; ++k)
        out->texels.push_back((uint32_t)rgba[4 * k] | (uint32_t)rgba[4 * k + 1] << 8 |
                              (uint32_t)rgba[4 * k + 2] << 16 | (uint32_t)rgba[4 * k + 3] << 24);
    }
    TextureDesc d;
    d.first = (uint32_t)image_first[(size_t)src];
    d.width = image_w[(size_t)src];
    d.height = image_h[(size_t)src];
    const int si = tx["sampler"].as_int(-1);
    if (si >= 0) {
      const JsonValue& sm = doc["samplers"][(size_t)si];
      auto mode_of = [](int gl) {
        return gl == 33071 ? kAddressClamp : gl == 33648 ? kAddressMirror : kAddressWrap;
      };
      d.wrap_s = mode_of(sm["wrapS"].as_int(10497));
      d.wrap_t = mode_of(sm["wrapT"].as_int(10497));
      d.filter = sm["minFilter"].as_int(-1) == 9728 ? kFilterPoint : kFilterLinear;
    }
    texture_slot[(size_t)t] = (int)out->textures.size();
    out->textures.push_back(d);
    return texture_slot[(size_t)t];
  }

  // Texture coordinate of a vertex: float, or normalised unsigned byte/short.
  static float get_texcoord(const Accessor& a, size_t i, int c) {
    const uint8_t* p = a.base + a.stride * i + (size_t)component_bytes(a.ctype) * c;
    switch (a.ctype) {
      case 5121: return (float)(*p) / 255.0f;
      case 5123: { uint16_t v; memcpy(&v, p, 2); return (float)v / 65535.0f; }
      default: { float f; memcpy(&f, p, 4); return f; }
    }
  }

  Mat4 node_matrix(const JsonValue& node) {
    Mat4 T = Mat4::identity(), R = Mat4::identity(), S = Mat4::identity(), M = Mat4::identity();
    const JsonValue& t = node["translation"];
    if (t.size() == 3) {
      T.m[3] = (float)t[0].num; T.m[7] = (float)t[1].num; T.m[11] = (float)t[2].num;
    }
    const JsonValue& r = node["rotation"];
    if (r.size() == 4) R = quat_matrix((float)r[3].num, (float)r[0].num, (float)r[1].num, (float)r[2].num);
    const JsonValue& s = node["scale"];
    if (s.size() == 3) {
      S.m[0] = (float)s[0].num; S.m[5] = (float)s[1].num; S.m[10] = (float)s[2].num;
    }
    const JsonValue& m = node["matrix"];
    if (m.size() == 16) {
      // glTF stores column-major
      for (int col = 0; col < 4; ++col)
        for (int row = 0; row < 4; ++row) M.m[4 * row + col] = (float)m[(size_t)(4 * col + row)].num;
    }
    return mul(mul(mul(M, T), R), S);
  }

  void transform_point(const Mat4& x, bool ident, const float in[3], float outp[3]) {
    if (ident) {
      outp[0] = in[0]; outp[1] = in[1]; outp[2] = in[2];
      return;
    }
    const float v[4] = {in[0], in[1], in[2], 1.0f};
    float o[4];
    apply(x, v, o);
    outp[0] = o[0]; outp[1] = o[1]; outp[2] = o[2];
  }

  // World box of an object-space box: the box around its 8 transformed
  // corners (sutil Aabb::transform, used by MulticamScene.cpp:388-389 and
  // hitscanprocessing.cpp:101-105).
  void transform_box(const Mat4& xf, const float lo[3], const float hi[3], float wlo[3], float whi[3]) {
    for (int a = 0; a < 3; ++a) {
      wlo[a] = INFINITY;
      whi[a] = -INFINITY;
    }
    for (int c = 0; c < 8; ++c) {
      const float v[4] = {(c & 4) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 1) ? hi[2] : lo[2], 1.0f};
      float o[4];
      apply(xf, v, o);
      for (int a = 0; a < 3; ++a) {
        wlo[a] = std::min(wlo[a], o[a]);
        whi[a] = std::max(whi[a], o[a]);
      }
    }
  }

  bool camera(const JsonValue& cam, const Mat4& xf) {
    CameraDesc cd;
    cd.name = cam["name"].as_string();
    const float vx[4] = {1, 0, 0, 0}, vy[4] = {0, 1, 0, 0}, vz[4] = {0, 0, -1, 0}, vp[4] = {0, 0, 0, 1};
    float o[4];
    apply(xf, vx, o); cd.pose.x[0] = o[0]; cd.pose.x[1] = o[1]; cd.pose.x[2] = o[2];
    apply(xf, vy, o); cd.pose.y[0] = o[0]; cd.pose.y[1] = o[1]; cd.pose.y[2] = o[2];
    apply(xf, vz, o); cd.pose.z[0] = o[0]; cd.pose.z[1] = o[1]; cd.pose.z[2] = o[2];
    apply(xf, vp, o); cd.pose.pos[0] = o[0]; cd.pose.pos[1] = o[1]; cd.pose.pos[2] = o[2];
    cd.yfov_deg = (float)cam["perspective"]["yfov"].as_number(1.0) * 180.0f / (float)M_PI;
    const JsonValue& extras = cam["extras"];
    if (cam["type"].as_string() == "orthographic") {
      cd.kind = kOrthographic;
      cd.xmag = (float)cam["orthographic"]["xmag"].as_number(1.0);
      cd.ymag = (float)cam["orthographic"]["ymag"].as_number(1.0);
    } else if (extras_true(extras, "panoramic")) {
      cd.kind = kPanoramic;
    } else if (extras_true(extras, "compound-eye")) {
      cd.kind = kCompound;
      const std::string path = extras["compound-structure"].as_string();
      cd.projection = extras["compound-projection"].as_string();
      if (path.empty()) {
        out->warnings.push_back("camera '" + cd.name + "': eye data path empty, camera skipped");
        return true;
      }
      if (cd.projection.empty()) {
        out->warnings.push_back("camera '" + cd.name + "': projection shader empty, camera skipped");
        return true;
      }
      std::string e1, e2;
      if (read_eye_file(path, &cd.ommatidia, &e1)) {
        cd.eye_path = path;
      } else if (read_eye_file(dir + path, &cd.ommatidia, &e2)) {
        cd.eye_path = dir + path;
      } else {
        out->warnings.push_back("camera '" + cd.name + "': " + e2 + ", camera skipped");
        return true;
      }
      if (cd.ommatidia.empty()) {
        out->warnings.push_back("camera '" + cd.name + "': zero ommatidia loaded, camera skipped");
        return true;
      }
    } else {
      cd.kind = kPerspective;
    }
    out->cameras.push_back(std::move(cd));
    return true;
  }

  bool mesh(const JsonValue& m, const Mat4& xf) {
    const bool ident = xf.is_identity();
    const std::string name = m["name"].as_string();
    if (extras_true(m["extras"], "hitbox")) {
      HitboxMesh hb;
      hb.name = name;
      float olo[3] = {INFINITY, INFINITY, INFINITY}, ohi[3] = {-INFINITY, -INFINITY, -INFINITY};
      const JsonValue& prims = m["primitives"];
      for (size_t pi = 0; pi < prims.size(); ++pi) {
        const JsonValue& prim = prims[pi];
        if (prim["mode"].as_int(4) != 4) continue;
        Accessor pa, ia;
        if (!accessor(prim["attributes"]["POSITION"].as_int(-1), &pa)) return false;
        if (pa.ctype != 5126 || pa.ncomp != 3) {
          err = "hitbox mesh '" + name + "': POSITION must be float VEC3";
          return false;
        }
        const int ii = prim["indices"].as_int(-1);
        if (ii >= 0) {
          if (!accessor(ii, &ia)) return false;
          if (ia.ncomp != 1 || (ia.ctype != 5121 && ia.ctype != 5123 && ia.ctype != 5125)) {
            err = "hitbox mesh '" + name + "': indices must be unsigned scalars";
            return false;
          }
        }
        const size_t ntri = ii >= 0 ? ia.count / 3 : pa.count / 3;
        for (size_t t = 0; t < ntri; ++t)
          for (int v = 0; v < 3; ++v) {
            const size_t vi = ii >= 0 ? get_uint(ia, 3 * t + v, 0) : 3 * t + v;
            if (vi >= pa.count) { err = "hitbox mesh '" + name + "': index out of range"; return false; }
            const float p[3] = {get_float(pa, vi, 0), get_float(pa, vi, 1), get_float(pa, vi, 2)};
            float w[3];
            transform_point(xf, ident, p, w);
            for (int a = 0; a < 3; ++a) {
              hb.tris.push_back(w[a]);
              olo[a] = std::min(olo[a], p[a]);
              ohi[a] = std::max(ohi[a], p[a]);
            }
          }
      }
      // object-space box of the vertices, then its transformed corners
      if (olo[0] <= ohi[0]) {
        transform_box(xf, olo, ohi, hb.lo, hb.hi);
      } else {
        for (int a = 0; a < 3; ++a) hb.lo[a] = hb.hi[a] = 0.0f;
      }
      out->hitboxes.push_back(std::move(hb));
      return true;
    }
    const JsonValue& prims = m["primitives"];
    for (size_t pi = 0; pi < prims.size(); ++pi) {
      const JsonValue& prim = prims[pi];
      if (prim["mode"].as_int(4) != 4) {
        out->warnings.push_back("mesh '" + name + "': non-triangle primitive skipped");
        continue;
      }
      const JsonValue& attrs = prim["attributes"];
      Accessor pa, ia, ca;
      if (!accessor(attrs["POSITION"].as_int(-1), &pa)) return false;
      if (pa.ctype != 5126 || pa.ncomp != 3) {
        err = "mesh '" + name + "': POSITION must be float VEC3";
        return false;
      }
      const int ii = prim["indices"].as_int(-1);
      if (ii >= 0) {
        if (!accessor(ii, &ia)) return false;
        if (ia.ncomp != 1 || (ia.ctype != 5121 && ia.ctype != 5123 && ia.ctype != 5125)) {
          err = "mesh '" + name + "': indices must be unsigned scalars";
          return false;
        }
      }
      // Vertex colours are used only when COLOR_0 is VEC4 (MulticamScene.cpp:415-505).
      bool use_vc = false;
      const int ci = attrs["COLOR_0"].as_int(-1);
      if (ci >= 0) {
        if (!accessor(ci, &ca)) return false;
        if (ca.ncomp == 4 && (ca.ctype == 5126 || ca.ctype == 5123 || ca.ctype == 5121)) {
          use_vc = true;
          if (ca.count < pa.count) {
            err = "mesh '" + name + "': COLOR_0 shorter than POSITION";
            return false;
          }
        } else {
          out->warnings.push_back("mesh '" + name + "': COLOR_0 is not VEC4 f32/u16/u8, vertex colours ignored");
        }
      }
      // Colour source in the reference's order (shaders.cu:788-806): vertex
      // colours, else the base-colour texture, else the base colour factor.
      float base[3] = {1.0f, 1.0f, 1.0f};
      int tex = -1;
      const int mi = prim["material"].as_int(-1);
      if (mi >= 0) {
        const JsonValue& mat = doc["materials"][(size_t)mi];
        const JsonValue& pbr = mat["pbrMetallicRoughness"];
        const JsonValue& bcf = pbr["baseColorFactor"];
        if (bcf.size() >= 3)
          for (int a = 0; a < 3; ++a) base[a] = (float)bcf[(size_t)a].num;
        if (!use_vc && pbr["baseColorTexture"].is_object()) {
          tex = texture(pbr["baseColorTexture"]["index"].as_int(-1));
          if (tex < 0) out->warnings.push_back("mesh '" + name + "': base colour texture unusable; using baseColorFactor");
        }
      }
      Accessor ta;
      bool has_uv = false;
      if (tex >= 0) {
        const int ti = attrs["TEXCOORD_0"].as_int(-1);
        if (ti >= 0) {
          if (!accessor(ti, &ta)) return false;
          if (ta.ncomp != 2 || (ta.ctype != 5126 && ta.ctype != 5121 && ta.ctype != 5123) || ta.count < pa.count) {
            err = "mesh '" + name + "': TEXCOORD_0 must be VEC2 float / normalised u8 / u16 covering POSITION";
            return false;
          }
          has_uv = true;
        }
      }
      const uint32_t mode_word = use_vc ? (uint32_t)kVertexColour
                                 : tex >= 0 ? ((uint32_t)kTextureColour | (uint32_t)tex << 2)
                                            : (uint32_t)kBaseColour;
      const size_t ntri = ii >= 0 ? ia.count / 3 : pa.count / 3;
      out->mesh_names.push_back(name);
      out->mesh_first_tri.push_back((uint32_t)out->mode.size());
      std::array<float, 6> mb = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
      out->pos.reserve(out->pos.size() + 9 * ntri);
      out->col.reserve(out->col.size() + 9 * ntri);
      out->mode.reserve(out->mode.size() + ntri);
      for (size_t t = 0; t < ntri; ++t) {
        for (int v = 0; v < 3; ++v) {
          const size_t vi = ii >= 0 ? get_uint(ia, 3 * t + v, 0) : 3 * t + v;
          if (vi >= pa.count) {
            err = "mesh '" + name + "': index out of range";
            return false;
          }
          const float p[3] = {get_float(pa, vi, 0), get_float(pa, vi, 1), get_float(pa, vi, 2)};
          float w[3];
          transform_point(xf, ident, p, w);
          out->pos.push_back(w[0]);
          out->pos.push_back(w[1]);
          out->pos.push_back(w[2]);
          for (int a = 0; a < 3; ++a) {
            mb[(size_t)a] = std::min(mb[(size_t)a], w[a]);
            mb[(size_t)a + 3] = std::max(mb[(size_t)a + 3], w[a]);
          }
          if (use_vc) {
            out->col.push_back(get_colour(ca, vi, 0));
            out->col.push_back(get_colour(ca, vi, 1));
            out->col.push_back(get_colour(ca, vi, 2));
          } else if (tex >= 0) {
            // two texture coordinates per vertex, then 3 unused slots
            if (has_uv) {
              out->col.push_back(get_texcoord(ta, vi, 0));
              out->col.push_back(get_texcoord(ta, vi, 1));
            } else {
              out->col.push_back(v == 1 ? 1.0f : 0.0f);
              out->col.push_back(v == 2 ? 1.0f : 0.0f);
            }
          } else {
            out->col.push_back(base[0]);
            out->col.push_back(base[1]);
            out->col.push_back(base[2]);
          }
        }
        if (tex >= 0 && !use_vc)
          for (int z = 0; z < 3; ++z) out->col.push_back(0.0f);
        out->mode.push_back(mode_word);
      }
      // World box from the POSITION accessor's declared min/max when present
      // (MulticamScene.cpp:377-389), else from the vertices.
      const JsonValue& pacc = doc["accessors"][(size_t)attrs["POSITION"].as_int(-1)];
      if (pacc["min"].size() >= 3 && pacc["max"].size() >= 3) {
        float olo[3], ohi[3], wlo[3], whi[3];
        for (int a = 0; a < 3; ++a) {
          olo[a] = (float)pacc["min"][(size_t)a].num;
          ohi[a] = (float)pacc["max"][(size_t)a].num;
        }
        transform_box(xf, olo, ohi, wlo, whi);
        for (int a = 0; a < 3; ++a) {
          mb[(size_t)a] = wlo[a];
          mb[(size_t)a + 3] = whi[a];
        }
      }
      out->mesh_bounds.push_back(mb);
    }
    return true;
  }

  bool node(int idx, const Mat4& parent, int depth) {
    if (depth > 64) {
      err = "node hierarchy too deep (cycle?)";
      return false;
    }
    const JsonValue& n = doc["nodes"][(size_t)idx];
    const Mat4 xf = mul(parent, node_matrix(n));
    // A camera node, a mesh node, or else recurse (MulticamScene.cpp:207-525).
    const int cam = n["camera"].as_int(-1);
    const int mesh_idx = n["mesh"].as_int(-1);
    if (cam >= 0) return camera(doc["cameras"][(size_t)cam], xf);
    if (mesh_idx >= 0) return mesh(doc["meshes"][(size_t)mesh_idx], xf);
    const JsonValue& ch = n["children"];
    for (size_t c = 0; c < ch.size(); ++c)
      if (!node(ch[c].as_int(0), xf, depth + 1)) return false;
    return true;
  }

  bool run(const std::string& path) {
    std::string text;
    if (!read_file(path, &text)) {
      err = "cannot open glTF file '" + path + "'";
      return false;
    }
    if (!json_parse(text, &doc, &err)) return false;
    const size_t slash = path.find_last_of("/\\");
    dir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
    out->background_shader = "__miss__default_background";
    const JsonValue& scenes = doc["scenes"];
    for (size_t i = 0; i < scenes.size(); ++i) {
      const std::string bg = scenes[i]["extras"]["background-shader"].as_string();
      if (!bg.empty()) out->background_shader = "__miss__" + bg;
    }
    if (out->background_shader == "__miss__simple_sky") {
      out->sky_mode = kSkySimple;
    } else {
      if (out->background_shader != "__miss__default_background")
        out->warnings.push_back("unknown background shader '" + out->background_shader +
                                "', using default_background");
      out->sky_mode = kSkyDefault;
    }
    if (!load_buffers()) return false;
    const JsonValue& nodes = doc["nodes"];
    std::vector<char> is_root(nodes.size(), 1);
    for (size_t i = 0; i < nodes.size(); ++i) {
      const JsonValue& ch = nodes[i]["children"];
      for (size_t c = 0; c < ch.size(); ++c) {
        const int k = ch[c].as_int(-1);
        if (k < 0 || (size_t)k >= nodes.size()) {
          err = "node " + std::to_string(i) + ": bad child index";
          return false;
        }
        is_root[(size_t)k] = 0;
      }
    }
    for (size_t i = 0; i < nodes.size(); ++i)
      if (is_root[i] && !node((int)i, Mat4::identity(), 0)) return false;
    return true;
  }
};

}  // namespace

bool load_gltf(const std::string& path, HostScene* out, std::string* err) {
  *out = HostScene();
  Loader L;
  L.out = out;
  if (!L.run(path)) {
    if (err) *err = "loadGlTFscene: " + L.err;
    *out = HostScene();
    return false;
  }
  return true;
}

bool read_eye_file(const std::string& path, std::vector<Ommatidium>* out, std::string* err) {
  std::ifstream f(path);
  if (!f.is_open()) {
    if (err) *err = "unable to open eye file '" + path + "'";
    return false;
  }
  out->clear();
  std::string line;
  size_t lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    bool blank = true;
    for (char c : line)
      if (!std::isspace((unsigned char)c)) blank = false;
    if (blank) continue;
    const std::vector<std::string> tok = split_on_space(line);
    std::vector<float> v;
    for (const std::string& t : tok) {
      if (t.empty()) continue;
      char* e = nullptr;
      const float x = strtof(t.c_str(), &e);
      if (e == t.c_str()) {
        if (err) *err = path + ":" + std::to_string(lineno) + ": not a number";
        return false;
      }
      v.push_back(x);
    }
    if (v.size() < 8) {
      if (err) *err = path + ":" + std::to_string(lineno) + ": expected 8 values";
      return false;
    }
    Ommatidium o;
    o.px = v[0]; o.py = v[1]; o.pz = v[2];
    o.dx = v[3]; o.dy = v[4]; o.dz = v[5];
    o.acceptance = v[6];
    o.focal_offset = v[7];
    out->push_back(o);
  }
  return true;
}

void pose_reset(Pose* p) {
  for (int a = 0; a < 3; ++a) {
    p->pos[a] = 0.0f;
    p->x[a] = a == 0 ? 1.0f : 0.0f;
    p->y[a] = a == 1 ? 1.0f : 0.0f;
    p->z[a] = a == 2 ? 1.0f : 0.0f;
  }
}

// rotatePoint with a normalised axis, DataRecordCamera.h:83-87.
static V3 rotate_normed(V3 p, float angle, V3 axis) {
  const V3 k = normalize(axis);
  float s, c;
  det_sincos(angle, &s, &c);
  return (c * p + s * cross(k, p)) + ((1.0f - c) * dot(k, p)) * k;
}

void pose_rotate_around(Pose* p, float angle, float ax, float ay, float az) {
  const V3 axis = v3(ax, ay, az);
  V3 X = rotate_normed(v3(p->x[0], p->x[1], p->x[2]), angle, axis);
  V3 Y = rotate_normed(v3(p->y[0], p->y[1], p->y[2]), angle, axis);
  V3 Z = rotate_normed(v3(p->z[0], p->z[1], p->z[2]), angle, axis);
  p->x[0] = X.x; p->x[1] = X.y; p->x[2] = X.z;
  p->y[0] = Y.x; p->y[1] = Y.y; p->y[2] = Y.z;
  p->z[0] = Z.x; p->z[1] = Z.y; p->z[2] = Z.z;
}

void pose_move(Pose* p, float x, float y, float z) {
  p->pos[0] = p->pos[0] + x;
  p->pos[1] = p->pos[1] + y;
  p->pos[2] = p->pos[2] + z;
}

}  // namespace eye
