// Per-ray device/host primitives of the compound-eye renderer: counter-based
// random numbers, Gaussian-cone ray generation, the watertight ray/triangle
// test, the robust slab test, vertex-colour shading and the sky miss colours.
//
// Everything is written against plain floats so that the same inline code is
// compiled for gfx950 (hipcc) and, for host-side helpers, for x86-64.  With
// -ffp-contract=off and the deterministic math in detmath.h every function
// returns identical bits on both sides.
#pragma once

#include "detmath.h"

namespace eye {

struct V3 {
  float x, y, z;
};

EYE_HD V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
EYE_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
EYE_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
EYE_HD V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
EYE_HD V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
EYE_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
EYE_HD V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// normalize as v * (1/sqrt(v.v)), the reference's formulation (sutil/vec_math.h:541-545)
EYE_HD V3 normalize(V3 v) {
  const float inv = 1.0f / sqrtf(dot(v, v));
  return v * inv;
}

// One ommatidium, 32 bytes: the reference's Ommatidium / OmmatidiumPacket
// layout (libEyeRenderer3/cameras/CompoundEyeDataTypes.h:22-28).
struct Ommatidium {
  float px, py, pz;
  float dx, dy, dz;
  float acceptance;   // radians, full width at half maximum of the cone
  float focal_offset; // tmin of the sample rays
};

// Eye pose: position plus the three local axes (GenericCameraDataTypes.h:17-42).
struct Pose {
  float pos[3];
  float x[3];
  float y[3];
  float z[3];
};

struct Ray {
  float ox, oy, oz, tmin;
  float dx, dy, dz, tmax;
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Multipliers and Weyl increments are the published ones.
// ---------------------------------------------------------------------------
struct U4 {
  uint32_t a, b, c, d;
};

EYE_HD uint32_t mulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(a, b);
#else
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
#endif
}

EYE_HD U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = mulhi32(M0, ctr.a), lo0 = M0 * ctr.a;
    const uint32_t hi1 = mulhi32(M1, ctr.c), lo1 = M1 * ctr.c;
    U4 n;
    n.a = hi1 ^ ctr.b ^ k0;
    n.b = lo1;
    n.c = hi0 ^ ctr.d ^ k1;
    n.d = lo0;
    ctr = n;
    k0 += W0;
    k1 += W1;
  }
  return ctr;
}

// Uniform in (0, 1] from the top 24 bits (the reference draws its azimuth from
// curand_uniform, whose range is (0, 1]; the same convention is kept).
EYE_HD float u01_open0(uint32_t r) { return (float)((r >> 8) + 1u) * 5.9604644775390625e-8f; }
// Uniform in [0, 1).
EYE_HD float u01_open1(uint32_t r) { return (float)(r >> 8) * 5.9604644775390625e-8f; }

// Standard normal from two uniforms, Box-Muller (cosine branch).
EYE_HD float box_muller(uint32_t r0, uint32_t r1) {
  const float u1 = u01_open0(r0);
  const float u2 = u01_open1(r1);
  const float rad = sqrtf(-2.0f * det_log(u1));
  return rad * det_cos(6.28318530717958648f * u2);
}

// FWHM / standard deviation of a Gaussian, libEyeRenderer3/shaders.cu:53.
constexpr float kFwhmSdRatio = 2.35482004503094938f;

// Rodrigues rotation of `p` by `angle` about `k` (k not normalised here, as in
// libEyeRenderer3/shaders.cu:648-651).
EYE_HD V3 rotate_point(V3 p, float angle, V3 k) {
  float s, c;
  det_sincos(angle, &s, &c);
  return (c * p + s * cross(k, p)) + ((1.0f - c) * dot(k, p)) * k;
}

// libEyeRenderer3/shaders.cu:652-662: splay the axis away from itself by
// `splay` about a perpendicular, then spin it by `spin` about the axis.
EYE_HD V3 offset_ray(float spin, float splay, V3 axis) {
  V3 perp = cross(v3(0.0f, 1.0f, 0.0f), axis);
  perp = (perp.x + perp.y + perp.z == 0.0f) ? v3(0.0f, 0.0f, 1.0f) : normalize(perp);
  const V3 splayed = rotate_point(axis, splay, perp);
  return rotate_point(splayed, spin, axis);
}

// Random stream of one sample: key (seed, eye), counter (ommatidium, sample,
// frame, pose).  One independent stream per (ommatidium, sample) that moves on
// once per frame, like the reference's per-id cuRAND states
// (libEyeRenderer3/shaders.cu:680-695), without any stored state.
EYE_HD U4 sample_randoms(uint32_t seed, uint32_t eye, uint32_t omm, uint32_t sample,
                         uint32_t frame, uint32_t pose) {
  U4 c;
  c.a = omm; c.b = sample; c.c = frame; c.d = pose;
  return philox4x32_10(c, seed, eye);
}

// Ray of one sample of one ommatidium in world space
// (libEyeRenderer3/shaders.cu:664-723).
EYE_HD Ray generate_ray(const Ommatidium& o, const Pose& p, U4 rnd) {
  const float sd = o.acceptance / kFwhmSdRatio;
  const float splay = box_muller(rnd.a, rnd.b) * sd;
  const float spin = u01_open0(rnd.c) * kPiF;
  const V3 axis = v3(o.dx, o.dy, o.dz);
  const V3 rdir = offset_ray(spin, splay, axis);
  const V3 rpos = v3(o.px, o.py, o.pz) - normalize(axis) * o.focal_offset;
  const V3 X = v3(p.x[0], p.x[1], p.x[2]);
  const V3 Y = v3(p.y[0], p.y[1], p.y[2]);
  const V3 Z = v3(p.z[0], p.z[1], p.z[2]);
  const V3 P = v3(p.pos[0], p.pos[1], p.pos[2]);
  const V3 org = ((P + X * rpos.x) + Y * rpos.y) + Z * rpos.z;
  const V3 dir = (X * rdir.x + Y * rdir.y) + Z * rdir.z;
  Ray r;
  r.ox = org.x; r.oy = org.y; r.oz = org.z; r.tmin = o.focal_offset;
  r.dx = dir.x; r.dy = dir.y; r.dz = dir.z; r.tmax = 1e16f;
  return r;
}

// ---------------------------------------------------------------------------
// Watertight ray/triangle intersection, Woop, Benthin, Wald, "Watertight
// Ray/Triangle Intersection", JCGT 2(1), 2013, in FP32 only (no FP64 fallback):
// an edge that evaluates to exactly zero counts as inside, so a ray through a
// shared edge reports both triangles and the (t, id) tie-break picks one.
// ---------------------------------------------------------------------------
struct RayPre {
  float ox, oy, oz;
  int kx, ky, kz;
  float sx, sy, sz;
  float tmin, tmax;
  // safe reciprocal direction for the slab test
  float ix, iy, iz;
  // bit a set when direction component a is (numerically) zero: the ray runs
  // parallel to that slab and is inside it iff lo <= origin <= hi
  int par;
};

EYE_HD float safe_rcp(float d) {
  const float tiny = 1e-30f;
  const float dd = fabsf(d) < tiny ? (d < 0.0f ? -tiny : tiny) : d;
  return 1.0f / dd;
}

EYE_HD float comp(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }

EYE_HD RayPre precompute_ray(const Ray& r) {
  RayPre p;
  p.ox = r.ox; p.oy = r.oy; p.oz = r.oz;
  const float ax = fabsf(r.dx), ay = fabsf(r.dy), az = fabsf(r.dz);
  int kz = 0;
  if (ay > ax) kz = 1;
  if (az > (kz == 0 ? ax : ay)) kz = 2;
  int kx = kz + 1; if (kx == 3) kx = 0;
  int ky = kx + 1; if (ky == 3) ky = 0;
  const float dkz = comp(r.dx, r.dy, r.dz, kz);
  if (dkz < 0.0f) { const int t = kx; kx = ky; ky = t; }
  p.kx = kx; p.ky = ky; p.kz = kz;
  p.sx = comp(r.dx, r.dy, r.dz, kx) / dkz;
  p.sy = comp(r.dx, r.dy, r.dz, ky) / dkz;
  p.sz = 1.0f / dkz;
  p.tmin = r.tmin; p.tmax = r.tmax;
  p.ix = safe_rcp(r.dx); p.iy = safe_rcp(r.dy); p.iz = safe_rcp(r.dz);
  p.par = (fabsf(r.dx) < 1e-30f ? 1 : 0) | (fabsf(r.dy) < 1e-30f ? 2 : 0) | (fabsf(r.dz) < 1e-30f ? 4 : 0);
  return p;
}

// Returns true and (t, V, W, det) when the ray hits the triangle with
// tmin <= t <= tmax; the barycentrics are b1 = V / det and b2 = W / det.
// Traversal keeps (V, W, det) of the closest hit and divides once at the end:
// the same operations on the same values as dividing at every hit.
EYE_HD bool intersect_triangle_raw(const RayPre& r, V3 v0, V3 v1, V3 v2, float* t_out, float* v_out,
                                   float* w_out, float* det_out) {
  const V3 A = v3(v0.x - r.ox, v0.y - r.oy, v0.z - r.oz);
  const V3 B = v3(v1.x - r.ox, v1.y - r.oy, v1.z - r.oz);
  const V3 C = v3(v2.x - r.ox, v2.y - r.oy, v2.z - r.oz);
  const float Akz = comp(A.x, A.y, A.z, r.kz), Bkz = comp(B.x, B.y, B.z, r.kz),
              Ckz = comp(C.x, C.y, C.z, r.kz);
  const float Ax = comp(A.x, A.y, A.z, r.kx) - r.sx * Akz;
  const float Ay = comp(A.x, A.y, A.z, r.ky) - r.sy * Akz;
  const float Bx = comp(B.x, B.y, B.z, r.kx) - r.sx * Bkz;
  const float By = comp(B.x, B.y, B.z, r.ky) - r.sy * Bkz;
  const float Cx = comp(C.x, C.y, C.z, r.kx) - r.sx * Ckz;
  const float Cy = comp(C.x, C.y, C.z, r.ky) - r.sy * Ckz;
  const float U = Cx * By - Cy * Bx;
  const float V = Ax * Cy - Ay * Cx;
  const float W = Bx * Ay - By * Ax;
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  const float det = U + V + W;
  if (det == 0.0f) return false;
  const float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
  const float T = U * Az + V * Bz + W * Cz;
  const float t = T / det;
  if (!(t >= r.tmin && t <= r.tmax)) return false;
  *t_out = t;
  *v_out = V;
  *w_out = W;
  *det_out = det;
  return true;
}

// Returns true and (t, b1, b2) when the ray hits the triangle with
// tmin <= t <= tmax.  b1, b2 weight vertices 1 and 2 (OptiX convention).
EYE_HD bool intersect_triangle(const RayPre& r, V3 v0, V3 v1, V3 v2, float* t_out,
                               float* b1_out, float* b2_out) {
  float V, W, det;
  if (!intersect_triangle_raw(r, v0, v1, v2, t_out, &V, &W, &det)) return false;
  *b1_out = V / det;
  *b2_out = W / det;
  return true;
}

// Robust slab test (Ize, "Robust BVH Ray Traversal", JCGT 2(2), 2013): the far
// distance is widened by 1 + 2*gamma(3) so that rounding never culls a box
// that holds a hit.  Returns the entry distance, or +inf for a miss.
constexpr float kRobustFar = 1.00000048f;  // 1 + 4 ulp >= 1 + 2*gamma(3)
// Boxes are culled against the current closest hit only when they start
// beyond it by more than this relative slack: the box entry and the triangle
// distance are rounded differently, and the (t, id) tie-break needs every
// triangle at the closest distance to be tested, including at shared vertices.
constexpr float kCullSlack = 1.0000152587890625f;  // 1 + 2^-16

// Slab entry/exit along one axis; a parallel ray gets (-inf, inf) inside the
// slab (boundaries included) and an empty interval outside it.
template <bool kParallel>
EYE_HD void slab_axis(float lo, float hi, float o, float inv, bool par, float* t0, float* t1) {
  if (kParallel && par) {
    *t0 = lo <= o ? -INFINITY : INFINITY;
    *t1 = hi >= o ? INFINITY : -INFINITY;
  } else {
    *t0 = (lo - o) * inv;
    *t1 = (hi - o) * inv;
  }
}

template <bool kParallel = false>
EYE_HD float slab_enter(const RayPre& r, float lox, float hix, float loy, float hiy,
                        float loz, float hiz, float tfar_limit) {
  float tx0, tx1, ty0, ty1, tz0, tz1;
  slab_axis<kParallel>(lox, hix, r.ox, r.ix, (r.par & 1) != 0, &tx0, &tx1);
  slab_axis<kParallel>(loy, hiy, r.oy, r.iy, (r.par & 2) != 0, &ty0, &ty1);
  slab_axis<kParallel>(loz, hiz, r.oz, r.iz, (r.par & 4) != 0, &tz0, &tz1);
  const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), r.tmin));
  const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1)) * kRobustFar;
  return (tn <= fminf(tf, tfar_limit)) ? tn : INFINITY;
}

// ---------------------------------------------------------------------------
// Shading (lighting off, libEyeRenderer.cpp:98): closest hit returns
// linearize(barycentric vertex colour) or the material base colour
// (libEyeRenderer3/shaders.cu:779-811, cuda/LocalGeometry.h:106-156).
// ---------------------------------------------------------------------------
constexpr float kGammaLinearize = 2.2f;

enum ColourMode : uint32_t { kVertexColour = 0u, kBaseColour = 1u, kTextureColour = 2u };
// The mode word of a triangle: ColourMode in bits 0-1, texture index above.
EYE_HD uint32_t colour_mode(uint32_t word) { return word & 3u; }
EYE_HD uint32_t texture_index(uint32_t word) { return word >> 2; }

enum AddressMode : int { kAddressWrap = 0, kAddressClamp = 1, kAddressMirror = 2 };
enum FilterMode : int { kFilterPoint = 0, kFilterLinear = 1 };

EYE_HD V3 shade_hit(float b1, float b2, V3 c0, V3 c1, V3 c2, uint32_t mode) {
  if (colour_mode(mode) == kBaseColour) return c0;
  const float w0 = (1.0f - b1) - b2;
  const V3 c = (w0 * c0 + b1 * c1) + b2 * c2;
  return v3(det_pow(c.x, kGammaLinearize), det_pow(c.y, kGammaLinearize),
            det_pow(c.z, kGammaLinearize));
}

// A base-colour texture as the kernels see it: texels first .. first + w*h - 1
// of the scene's RGBA8 texel array.
struct TexView {
  int width, height, wrap_s, wrap_t, filter;
  uint32_t first;
};

// Texel index along one axis under an address mode: repeat, clamp to the
// edge, or mirrored repeat (the three modes the reference maps glTF samplers
// to, MulticamScene.cpp:613-621).
EYE_HD int texel_index(int i, int n, int mode) {
  if (mode == kAddressClamp) return i < 0 ? 0 : (i >= n ? n - 1 : i);
  if (mode == kAddressMirror) {
    const int period = 2 * n;
    int m = i % period;
    if (m < 0) m += period;
    return m < n ? m : period - 1 - m;
  }
  int m = i % n;
  if (m < 0) m += n;
  return m;
}

// Normalised coordinate folded into one period first, so that texel indices
// stay small: [0, 1) for repeat, [0, 2) for mirrored repeat, [0, 1] for clamp.
EYE_HD float fold_coord(float u, int mode) {
  if (!(u == u)) return 0.0f;
  if (mode == kAddressClamp) return fminf(fmaxf(u, 0.0f), 1.0f);
  if (mode == kAddressMirror) return u - 2.0f * floorf(u * 0.5f);
  return u - floorf(u);
}

EYE_HD V3 texel_rgb(const uint32_t* texels, const TexView& t, int i, int j) {
  const uint32_t v = texels[t.first + (uint32_t)j * (uint32_t)t.width + (uint32_t)i];
  return v3((float)(v & 0xffu) / 255.0f, (float)((v >> 8) & 0xffu) / 255.0f, (float)((v >> 16) & 0xffu) / 255.0f);
}

// tex2D with normalised coordinates and normalised-float reads: point
// sampling takes texel floor(u W); linear filtering blends the four texels
// around (u W - 0.5, v H - 0.5) with weights quantised to 1/256 (the 8-bit
// fraction of the texture unit's filter weights, as CUDA documents it).
EYE_HD V3 sample_texture(const uint32_t* texels, const TexView& t, float u, float v) {
  const float x = fold_coord(u, t.wrap_s) * (float)t.width;
  const float y = fold_coord(v, t.wrap_t) * (float)t.height;
  if (t.filter == kFilterPoint) {
    const int i = texel_index((int)floorf(x), t.width, t.wrap_s);
    const int j = texel_index((int)floorf(y), t.height, t.wrap_t);
    return texel_rgb(texels, t, i, j);
  }
  const float xb = x - 0.5f, yb = y - 0.5f;
  const float fx = floorf(xb), fy = floorf(yb);
  const float a = rintf((xb - fx) * 256.0f) * (1.0f / 256.0f);
  const float b = rintf((yb - fy) * 256.0f) * (1.0f / 256.0f);
  const int i0 = texel_index((int)fx, t.width, t.wrap_s), i1 = texel_index((int)fx + 1, t.width, t.wrap_s);
  const int j0 = texel_index((int)fy, t.height, t.wrap_t), j1 = texel_index((int)fy + 1, t.height, t.wrap_t);
  const V3 t00 = texel_rgb(texels, t, i0, j0), t10 = texel_rgb(texels, t, i1, j0);
  const V3 t01 = texel_rgb(texels, t, i0, j1), t11 = texel_rgb(texels, t, i1, j1);
  const V3 top = t00 * (1.0f - a) + t10 * a;
  const V3 bottom = t01 * (1.0f - a) + t11 * a;
  return top * (1.0f - b) + bottom * b;
}

// Textured hit: texture coordinates interpolated like vertex colours
// (LocalGeometry.h:92-103; without TEXCOORD_0 the loader stores (0,0),
// (1,0), (0,1), which gives the barycentrics), then linearize of the texel
// colour (shaders.cu:801-805).  uv holds u0 v0 u1 v1 u2 v2.
EYE_HD V3 shade_texture(float b1, float b2, const float uv[6], const uint32_t* texels, const TexView& t) {
  const float w0 = (1.0f - b1) - b2;
  const float u = (w0 * uv[0] + b1 * uv[2]) + b2 * uv[4];
  const float v = (w0 * uv[1] + b1 * uv[3]) + b2 * uv[5];
  const V3 c = sample_texture(texels, t, u, v);
  return v3(det_pow(c.x, kGammaLinearize), det_pow(c.y, kGammaLinearize), det_pow(c.z, kGammaLinearize));
}

// __miss__simple_sky, libEyeRenderer3/shaders.cu:749-756.
EYE_HD V3 sky_simple(float dx, float dy, float dz) {
  const V3 d = normalize(v3(dx, dy, dz));
  const float mix = fminf(fmaxf(0.0f, (det_asin(d.y) * 2.0f) / kPiF), 1.0f);
  const V3 upper = v3(1.0f / 255.0f, 31.0f / 255.0f, 117.0f / 255.0f);
  const V3 lower = v3(143.0f / 255.0f, 179.0f / 255.0f, 203.0f / 255.0f) * 0.8f;
  return lower * (1.0f - mix) + upper * mix;
}

// __miss__default_background, libEyeRenderer3/shaders.cu:740-747.
EYE_HD V3 sky_default(float dx, float dy, float dz) {
  const V3 d = normalize(v3(dx, dy, dz));
  const float border = 0.01f;
  if (fabsf(d.x) < border || fabsf(d.y) < border || fabsf(d.z) < border) return v3(0.0f, 0.0f, 0.0f);
  return v3((det_atan2(d.z, d.x) + kPiF) / (kPiF * 2.0f), (det_asin(d.y) + kPiF / 2.0f) / kPiF, 0.0f);
}

enum SkyMode : int { kSkyDefault = 0, kSkySimple = 1 };

EYE_HD V3 sky_colour(int mode, float dx, float dy, float dz) {
  return mode == kSkySimple ? sky_simple(dx, dy, dz) : sky_default(dx, dy, dz);
}

// make_color (libEyeRenderer3/shaders.cu:180-189): gamma 1/2.2, truncation.
constexpr float kGammaEncode = 0.454545438289642333984375f;  // (float)(1.0 / (double)2.2f)

EYE_HD uint8_t encode_channel(float c) {
  const float cc = fminf(fmaxf(c, 0.0f), 1.0f);
  return (uint8_t)(det_pow(cc, kGammaEncode) * 255.0f);
}

// ---- regular cameras (shaders.cu:198-333) ----
enum CameraRayKind : int { kCameraPinhole = 0, kCameraOrthographic = 1, kCameraPanoramic = 2 };

struct CameraLaunch {
  Pose pose;
  int kind;                   // CameraRayKind
  float scale_x, scale_y, scale_z;  // pinhole: tan(fov/2) x, y and 1; orthographic: xmag, ymag
  float start_radius;         // panoramic
  int width, height;
};

// ---- spherical projections (shaders.cu:412-640) ----
enum SphericalVariant : int { kSphericalPosition = 0, kSphericalOrientation = 1, kSphericalSplit = 2 };

// Unit-sphere direction shown by pixel (x, y) of a width x height window:
// longitude pi/2 - pi d.x, latitude pi/2 d.y, d = 2 (x/w, y/h) - 1.  The
// split variant maps each half of the window onto the whole sphere.
EYE_HD V3 sphere_direction(int x, int y, int width, int height, bool split) {
  const float ux = (float)x / (float)width, uy = (float)y / (float)height;
  float dx, dy;
  if (split) {
    const float sx = ux * 2.0f;
    dx = (sx - (sx > 1.0f ? 1.0f : 0.0f)) * 2.0f - 1.0f;
    dy = uy * 2.0f - 1.0f;
  } else {
    dx = 2.0f * ux - 1.0f;
    dy = 2.0f * uy - 1.0f;
  }
  const float lon = dx * -kPiF + kPiF / 2.0f;
  const float lat = dy * (kPiF / 2.0f) + 0.0f;
  float slon, clon, slat, clat;
  det_sincos(lon, &slon, &clon);
  det_sincos(lat, &slat, &clat);
  return v3(clon * clat, slat, slon * clat);
}

// Primary ray of pixel (x, y) of a regular camera, d = 2 (x/w, y/h) - 1:
// pinhole: direction Z sz + d.x X sx + d.y Y sy from the position (not
// normalised); orthographic: direction Z from position + d.x X sx + d.y Y sy;
// panoramic: the unit direction of sphere_direction in the camera's axes,
// from position + direction * start radius.  tmin 0.01, tmax 1e16.
EYE_HD Ray camera_ray(const CameraLaunch& C, int x, int y) {
  const Pose& P = C.pose;
  const V3 X = v3(P.x[0], P.x[1], P.x[2]), Y = v3(P.y[0], P.y[1], P.y[2]), Z = v3(P.z[0], P.z[1], P.z[2]);
  const V3 O = v3(P.pos[0], P.pos[1], P.pos[2]);
  const float dx = 2.0f * ((float)x / (float)C.width) - 1.0f;
  const float dy = 2.0f * ((float)y / (float)C.height) - 1.0f;
  V3 o, d;
  if (C.kind == kCameraOrthographic) {
    d = Z;
    o = (O + (dx * X) * C.scale_x) + (dy * Y) * C.scale_y;
  } else if (C.kind == kCameraPanoramic) {
    const V3 s = sphere_direction(x, y, C.width, C.height, false);
    d = normalize((s.x * X + s.y * Y) + s.z * Z);
    o = O + d * C.start_radius;
  } else {
    d = (Z * C.scale_z + (dx * X) * C.scale_x) + (dy * Y) * C.scale_y;
    o = O;
  }
  Ray r;
  r.ox = o.x; r.oy = o.y; r.oz = o.z;
  r.tmin = 0.01f;
  r.dx = d.x; r.dy = d.y; r.dz = d.z;
  r.tmax = 1e16f;
  return r;
}

// Cosine of the angle between v and the unit-sphere point s (the reference
// compares acos of this value; the smallest angle is the largest cosine).
EYE_HD float cos_angle(V3 v, V3 s, float len_s) {
  const float len_v = sqrtf((v.x * v.x + v.y * v.y) + v.z * v.z);
  return ((v.x * s.x + v.y * s.y) + v.z * s.z) / (len_v * len_s);
}

}  // namespace eye
