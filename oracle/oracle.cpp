// CPU ORACLE — test infrastructure, not product code.
//
// Scalar C++ restatement of the compound-eye rendering path of the reference
// (BrainsOnBoard/compound-ray), used only by tests/, the smoke check and the
// CPU-baseline leg of bench.py, to check the GPU results against.  It
// re-derives every step from the reference sources and the documented design
// rules, independently of the GPU library:
//   ray generation   libEyeRenderer3/shaders.cu:648-723 (generateOffsetRay,
//                    rotatePoint, __raygen__ommatidium), FWHM ratio :53
//   random streams   one stream per (ommatidium, sample) advanced per frame,
//                    shaders.cu:680-695 — realised as Philox4x32-10 (Salmon et
//                    al., SC'11) keyed (seed, eye), counter (omm, sample,
//                    frame, pose); cuRAND's XORWOW is deliberately not
//                    reproduced
//   closest hit      OptiX traversal (closed) — restated as the exact
//                    lexicographic minimum of (t, triangle id) over all
//                    triangles with tmin <= t <= tmax, watertight test of
//                    Woop/Benthin/Wald (JCGT 2013) without FP64 fallback
//   shading          shaders.cu:779-811 + cuda/LocalGeometry.h:106-156
//                    (barycentric vertex colour, pow 2.2; base colour
//                    otherwise; lighting off, libEyeRenderer.cpp:98)
//   miss             shaders.cu:740-756 (default_background, simple_sky)
//   reduction        shaders.cu:730 + :341-347 (c * (1/spp), summed) in two
//                    orders: the reference's sequential order and the GPU's
//                    canonical wave order; an FP64-accumulated variant
//   readback         make_color, shaders.cu:180-189
//   acceleration     LBVH with the deterministic rules of the GPU builder
//                    (Morton codes from FP64-quantised centroids, stable
//                    (code, id) sort, Karras 2012 hierarchy, min/max refit,
//                    breadth-first collapse to WIDTH children by largest surface area)
//   learned sampler  FP64 MLP forward/backward for gradient checks
// Transcendentals use the same published series as the GPU library (Taylor,
// atanh-series log, Maclaurin asin) built from IEEE +,-,*,/,fma,sqrt only, so
// results are bit-identical to the GPU's where the arithmetic is specified.
//
// Compile: g++ -O2 -std=c++17 -ffp-contract=off -fPIC -shared -pthread.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// ----------------------------------------------------------------- math ----
constexpr float PI_F = 3.14159265358979323846f;
constexpr float PIO2_A = 1.57079637050628662109375f;
constexpr float PIO2_B = -4.371138828673793e-8f;
constexpr float PIO2_C = -1.7151245100058819e-15f;

uint32_t bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
float from_bits(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

void o_sincos(float x, float* so, float* co) {
  const float k = std::rint(x * 0.636619772367581343f);
  float r = std::fma(-k, PIO2_A, x);
  r = std::fma(-k, PIO2_B, r);
  r = std::fma(-k, PIO2_C, r);
  const float r2 = r * r;
  static const float S[5] = {-2.50521083854417188e-8f, 2.75573192239858907e-6f, -1.98412698412698413e-4f,
                             8.33333333333333333e-3f, -1.66666666666666667e-1f};
  static const float C[6] = {2.08767569878680990e-9f, -2.75573192239858907e-7f, 2.48015873015873016e-5f,
                             -1.38888888888888889e-3f, 4.16666666666666667e-2f, -0.5f};
  float ps = S[0];
  for (int i = 1; i < 5; ++i) ps = std::fma(ps, r2, S[i]);
  const float s = std::fma(ps * r2, r, r);
  float pc = C[0];
  for (int i = 1; i < 6; ++i) pc = std::fma(pc, r2, C[i]);
  const float c = std::fma(pc, r2, 1.0f);
  switch (((int)k) & 3) {
    case 0: *so = s; *co = c; break;
    case 1: *so = c; *co = -s; break;
    case 2: *so = -s; *co = -c; break;
    default: *so = -c; *co = s; break;
  }
}

float o_log(float x) {
  if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
  if (x == INFINITY) return INFINITY;
  uint32_t b = bits(x);
  int e = 0;
  if (b < 0x00800000u) { x = x * 8388608.0f; b = bits(x); e = -23; }
  e += (int)((b >> 23) & 0xffu) - 127;
  uint32_t mb = (b & 0x007fffffu) | 0x3f800000u;
  if (mb > 0x3fb504f3u) { mb -= 0x00800000u; e += 1; }
  const float m = from_bits(mb);
  const float s = (m - 1.0f) / (m + 1.0f);
  const float s2 = s * s;
  static const float P[5] = {9.09090909090909091e-2f, 1.11111111111111111e-1f, 1.42857142857142857e-1f, 2.0e-1f,
                             3.33333333333333333e-1f};
  float p = P[0];
  for (int i = 1; i < 5; ++i) p = std::fma(p, s2, P[i]);
  const float two_s = s + s;
  const float lnm = std::fma(two_s * s2, p, two_s);
  const float ef = (float)e;
  return std::fma(ef, 0.693145751953125f, std::fma(ef, 1.42860682030941723e-6f, lnm));
}

float o_exp2(float x) {
  if (x != x) return x;
  if (x < -160.0f) return 0.0f;
  if (x > 128.0f) return INFINITY;
  const float k = std::rint(x);
  const float f = x - k;
  static const float Q[10] = {1.01780860092396997e-7f, 1.32154867901443094e-6f, 1.52527338040598403e-5f,
                              1.54035303933816099e-4f, 1.33335581464284434e-3f, 9.61812910762847716e-3f,
                              5.55041086648215800e-2f, 2.40226506959100712e-1f, 6.93147180559945309e-1f, 1.0f};
  float p = Q[0];
  for (int i = 1; i < 10; ++i) p = std::fma(p, f, Q[i]);
  int ki = (int)k;
  if (ki < -126) { p = p * from_bits(0x00800000u); ki += 126; }
  else if (ki > 127) { p = p * from_bits(0x7f000000u); ki -= 127; }
  return p * from_bits((uint32_t)(ki + 127) << 23);
}

float o_pow(float x, float y) {
  if (!(x > 0.0f)) return 0.0f;
  if (x == 1.0f) return 1.0f;
  return o_exp2(y * (o_log(x) * 1.44269504088896341f));
}

float o_asin_small(float x) {
  const float x2 = x * x;
  static const float A[10] = {8.39033580961681500e-3f, 9.76160952919440038e-3f, 1.15518008961397050e-2f,
                              1.39648437500000000e-2f, 1.73527644230769231e-2f, 2.23721590909090909e-2f,
                              3.03819444444444444e-2f, 4.46428571428571429e-2f, 7.5e-2f, 1.66666666666666667e-1f};
  float p = A[0];
  for (int i = 1; i < 10; ++i) p = std::fma(p, x2, A[i]);
  return std::fma(x * x2, p, x);
}

float o_asin(float x) {
  const float ax = std::fabs(x);
  if (ax <= 0.5f) return o_asin_small(x);
  if (ax > 1.0f) return NAN;
  const float a = o_asin_small(std::sqrt((1.0f - ax) * 0.5f));
  const float v = std::fma(-2.0f, a, PIO2_A) + PIO2_B;
  return x < 0.0f ? -v : v;
}

float o_atan_unit(float t) {
  float off = 0.0f;
  if (t > 0.414213562373095049f) {
    t = (t - 1.0f) / (t + 1.0f);
    off = 0.785398163397448310f;
  }
  const float t2 = t * t;
  static const float T[10] = {4.76190476190476190e-2f, -5.26315789473684211e-2f, 5.88235294117647059e-2f,
                              -6.66666666666666667e-2f, 7.69230769230769231e-2f, -9.09090909090909091e-2f,
                              1.11111111111111111e-1f, -1.42857142857142857e-1f, 2.0e-1f, -3.33333333333333333e-1f};
  float p = T[0];
  for (int i = 1; i < 10; ++i) p = std::fma(p, t2, T[i]);
  return off + std::fma(t * t2, p, t);
}

float o_atan2(float y, float x) {
  if (x != x || y != y) return NAN;
  const float ax = std::fabs(x), ay = std::fabs(y);
  if (ax == 0.0f && ay == 0.0f) return 0.0f;
  float a = ay <= ax ? o_atan_unit(ay / ax) : PIO2_A - o_atan_unit(ax / ay);
  if (x < 0.0f) a = PI_F - a;
  return y < 0.0f ? -a : a;
}

// --------------------------------------------------------------- philox ----
void philox(const uint32_t in[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// --------------------------------------------------------------- vectors ----
struct F3 { float x, y, z; };
F3 f3(float x, float y, float z) { return F3{x, y, z}; }
F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
F3 scale(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
F3 cross(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
F3 normalize(F3 v) { return scale(v, 1.0f / std::sqrt(dot(v, v))); }  // sutil/vec_math.h:541-545

struct Omm { float p[3], d[3], acceptance, focal; };   // CompoundEyeDataTypes.h:22-28
struct PoseO { float pos[3], x[3], y[3], z[3]; };      // GenericCameraDataTypes.h:17-42
struct RayO { float o[3], tmin, d[3], tmax; };

// shaders.cu:648-651 (axis used as given)
F3 rotate_point(F3 p, float angle, F3 k) {
  float s, c;
  o_sincos(angle, &s, &c);
  return add(add(scale(p, c), scale(cross(k, p), s)), scale(k, (1.0f - c) * dot(k, p)));
}

RayO make_ray(const Omm& om, const PoseO& P, uint32_t seed, uint32_t eye, uint32_t omm, uint32_t sample,
              uint32_t frame, uint32_t pose) {
  const uint32_t ctr[4] = {omm, sample, frame, pose};
  uint32_t r[4];
  philox(ctr, seed, eye, r);
  const float u1 = (float)((r[0] >> 8) + 1u) * 5.9604644775390625e-8f;   // (0, 1]
  const float u2 = (float)(r[1] >> 8) * 5.9604644775390625e-8f;          // [0, 1)
  float s2, c2;
  o_sincos(6.28318530717958648f * u2, &s2, &c2);
  const float normal = std::sqrt(-2.0f * o_log(u1)) * c2;
  const float sd = om.acceptance / 2.35482004503094938f;   // FWHM_SD_RATIO, shaders.cu:53
  const float splay = normal * sd;                         // shaders.cu:690-691
  const float spin = ((float)((r[2] >> 8) + 1u) * 5.9604644775390625e-8f) * PI_F;  // shaders.cu:692
  const F3 axis = f3(om.d[0], om.d[1], om.d[2]);
  F3 perp = cross(f3(0.0f, 1.0f, 0.0f), axis);              // shaders.cu:655-657
  perp = (perp.x + perp.y + perp.z == 0.0f) ? f3(0.0f, 0.0f, 1.0f) : normalize(perp);
  const F3 dl = rotate_point(rotate_point(axis, splay, perp), spin, axis);   // :659-661
  const F3 pl = sub(f3(om.p[0], om.p[1], om.p[2]), scale(normalize(axis), om.focal));  // :701
  const F3 X = f3(P.x[0], P.x[1], P.x[2]), Y = f3(P.y[0], P.y[1], P.y[2]), Z = f3(P.z[0], P.z[1], P.z[2]);
  const F3 org = add(add(add(f3(P.pos[0], P.pos[1], P.pos[2]), scale(X, pl.x)), scale(Y, pl.y)), scale(Z, pl.z));
  const F3 dir = add(add(scale(X, dl.x), scale(Y, dl.y)), scale(Z, dl.z));   // :704-709
  RayO ray;
  ray.o[0] = org.x; ray.o[1] = org.y; ray.o[2] = org.z; ray.tmin = om.focal;   // :717-722
  ray.d[0] = dir.x; ray.d[1] = dir.y; ray.d[2] = dir.z; ray.tmax = 1e16f;
  return ray;
}

// ------------------------------------------------------------ intersect ----
struct Pre {
  float o[3];
  int kx, ky, kz;
  float sx, sy, sz, tmin, tmax;
  float inv[3];
  bool par[3];
};

Pre prepare(const RayO& r) {
  Pre p;
  for (int a = 0; a < 3; ++a) p.o[a] = r.o[a];
  const float ad[3] = {std::fabs(r.d[0]), std::fabs(r.d[1]), std::fabs(r.d[2])};
  int kz = 0;
  if (ad[1] > ad[0]) kz = 1;
  if (ad[2] > ad[kz]) kz = 2;
  int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
  if (r.d[kz] < 0.0f) std::swap(kx, ky);
  p.kx = kx; p.ky = ky; p.kz = kz;
  p.sx = r.d[kx] / r.d[kz];
  p.sy = r.d[ky] / r.d[kz];
  p.sz = 1.0f / r.d[kz];
  p.tmin = r.tmin;
  p.tmax = r.tmax;
  for (int a = 0; a < 3; ++a) {
    float d = r.d[a];
    p.par[a] = std::fabs(d) < 1e-30f;
    if (p.par[a]) d = d < 0.0f ? -1e-30f : 1e-30f;
    p.inv[a] = 1.0f / d;
  }
  return p;
}

bool hit_triangle(const Pre& r, const float* v, float* t_out, float* b1_out, float* b2_out) {
  float A[3], B[3], C[3];
  for (int a = 0; a < 3; ++a) {
    A[a] = v[a] - r.o[a];
    B[a] = v[3 + a] - r.o[a];
    C[a] = v[6 + a] - r.o[a];
  }
  const float Ax = A[r.kx] - r.sx * A[r.kz], Ay = A[r.ky] - r.sy * A[r.kz];
  const float Bx = B[r.kx] - r.sx * B[r.kz], By = B[r.ky] - r.sy * B[r.kz];
  const float Cx = C[r.kx] - r.sx * C[r.kz], Cy = C[r.ky] - r.sy * C[r.kz];
  const float U = Cx * By - Cy * Bx;
  const float V = Ax * Cy - Ay * Cx;
  const float W = Bx * Ay - By * Ax;
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  const float det = U + V + W;
  if (det == 0.0f) return false;
  const float T = U * (r.sz * A[r.kz]) + V * (r.sz * B[r.kz]) + W * (r.sz * C[r.kz]);
  const float t = T / det;
  if (!(t >= r.tmin && t <= r.tmax)) return false;
  *t_out = t;
  *b1_out = V / det;
  *b2_out = W / det;
  return true;
}

// ------------------------------------------------------------------ BVH ----
constexpr int32_t EMPTY = 0x7fffffff;
constexpr float EMPTY_BOX = 3.0e38f;

// Nodes have up to WIDTH children (4 or 8, the library's build width).
// Memory layout (32 * WIDTH bytes): WIDTH / 4 halves of 128 bytes with four
// slots each, every half holding its slots' boxes as arrays (lo x, hi x,
// lo y, hi y, lo z, hi z), the child words and four padding words; slot s is
// in half s / 4 at position s % 4.
constexpr int WIDTH = 4;
struct NodeHalf {
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  int32_t child[4];
  uint32_t pad[4];
};
struct Node {
  NodeHalf h[WIDTH / 4];
  const NodeHalf& of(int s) const { return h[s / 4]; }
  NodeHalf& of(int s) { return h[s / 4]; }
  int32_t child(int s) const { return h[s / 4].child[s % 4]; }
};
static_assert(sizeof(Node) == 32 * WIDTH, "node layout");

float mn(float a, float b) { return a < b ? a : b; }
float mx(float a, float b) { return a > b ? a : b; }

uint32_t spread10(uint32_t v) {
  uint32_t r = 0;
  for (int i = 0; i < 10; ++i) r |= ((v >> i) & 1u) << (3 * i);
  return r;
}

uint32_t quant10(float c, float lo, float hi) {
  const double ext = (double)hi - (double)lo;
  if (!(ext > 0.0)) return 0;
  double q = std::floor(((double)c - (double)lo) / ext * 1024.0);
  q = q < 0.0 ? 0.0 : (q > 1023.0 ? 1023.0 : q);
  return (uint32_t)q;
}

// Total order on floats with -0 < +0 (the GPU reduces centroid bounds this way).
bool key_less(float a, float b) {
  auto key = [](float f) {
    const uint32_t u = bits(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  };
  return key(a) < key(b);
}

struct Scene {
  int n = 0;
  int sky = 0;
  int leaf_max = 4;
  std::vector<float> pos;   // 9 per triangle, original order
  std::vector<float> col;   // 9 per triangle
  std::vector<uint32_t> mode;
  // base-colour textures: (width, height, wrap s, wrap t, filter, first texel) each
  std::vector<int32_t> tex;
  std::vector<uint32_t> texels;      // RGBA8, red in the low byte
  std::vector<uint32_t> codes, ids;  // sorted
  std::vector<Node> nodes;
  std::vector<float> stris;  // 9 per sorted triangle
  int stack_bound = 0;
  int levels = 0;
  float clo[3], chi[3];
};

int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }

int delta(const std::vector<uint32_t>& k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  if (k[i] != k[j]) return clz32(k[i] ^ k[j]);
  return 32 + clz32((uint32_t)i ^ (uint32_t)j);
}

void build(Scene& S) {
  const int n = S.n;
  S.nodes.clear();
  S.stack_bound = 0;
  S.levels = 0;
  if (n == 0) return;
  std::vector<float> tb(6 * (size_t)n), cen(3 * (size_t)n);
  for (int a = 0; a < 3; ++a) { S.clo[a] = 0; S.chi[a] = 0; }
  for (int i = 0; i < n; ++i) {
    const float* p = &S.pos[9 * (size_t)i];
    for (int a = 0; a < 3; ++a) {
      const float lo = mn(mn(p[a], p[3 + a]), p[6 + a]);
      const float hi = mx(mx(p[a], p[3 + a]), p[6 + a]);
      tb[6 * (size_t)i + a] = lo;
      tb[6 * (size_t)i + 3 + a] = hi;
      const float c = (lo + hi) * 0.5f;
      cen[3 * (size_t)i + a] = c;
      if (i == 0 || key_less(c, S.clo[a])) S.clo[a] = c;
      if (i == 0 || key_less(S.chi[a], c)) S.chi[a] = c;
    }
  }
  std::vector<std::pair<uint32_t, uint32_t>> kv((size_t)n);
  for (int i = 0; i < n; ++i) {
    const uint32_t code = (spread10(quant10(cen[3 * (size_t)i], S.clo[0], S.chi[0])) << 2) |
                          (spread10(quant10(cen[3 * (size_t)i + 1], S.clo[1], S.chi[1])) << 1) |
                          spread10(quant10(cen[3 * (size_t)i + 2], S.clo[2], S.chi[2]));
    kv[(size_t)i] = {code, (uint32_t)i};
  }
  std::sort(kv.begin(), kv.end());
  S.codes.resize((size_t)n);
  S.ids.resize((size_t)n);
  for (int i = 0; i < n; ++i) { S.codes[(size_t)i] = kv[(size_t)i].first; S.ids[(size_t)i] = kv[(size_t)i].second; }
  // binary tree, combined indexing: internal 0..n-2, leaf k at n-1+k
  const int nb = 2 * n - 1;
  std::vector<int> left((size_t)std::max(n - 1, 1)), right((size_t)std::max(n - 1, 1));
  std::vector<int> rfirst((size_t)nb), rlast((size_t)nb);
  std::vector<float> bnd(6 * (size_t)nb);
  for (int k = 0; k < n; ++k) {
    rfirst[(size_t)(n - 1 + k)] = k;
    rlast[(size_t)(n - 1 + k)] = k;
    for (int a = 0; a < 6; ++a) bnd[6 * (size_t)(n - 1 + k) + a] = tb[6 * (size_t)S.ids[(size_t)k] + a];
  }
  const std::vector<uint32_t>& K = S.codes;
  for (int i = 0; i < n - 1; ++i) {
    const int d = (delta(K, n, i, i + 1) - delta(K, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(K, n, i, i - d);
    int lmax = 2;
    while (delta(K, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
      if (delta(K, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dn = delta(K, n, i, j);
    int s = 0;
    for (int div = 2;; div *= 2) {
      const int t = (l + div - 1) / div;
      if (delta(K, n, i, i + (s + t) * d) > dn) s += t;
      if (t <= 1) break;
    }
    const int g = i + s * d + (d < 0 ? -1 : 0);
    const int lo = std::min(i, j), hi = std::max(i, j);
    left[(size_t)i] = lo == g ? n - 1 + g : g;
    right[(size_t)i] = hi == g + 1 ? n - 1 + g + 1 : g + 1;
    rfirst[(size_t)i] = lo;
    rlast[(size_t)i] = hi;
  }
  // refit: children of internal i have larger combined index only for leaves;
  // process internal nodes in post-order via an explicit stack
  if (n > 1) {
    std::vector<int> order;
    order.reserve((size_t)n);
    std::vector<int> st = {0};
    while (!st.empty()) {
      const int v = st.back();
      st.pop_back();
      order.push_back(v);
      if (left[(size_t)v] < n - 1) st.push_back(left[(size_t)v]);
      if (right[(size_t)v] < n - 1) st.push_back(right[(size_t)v]);
    }
    for (size_t q = order.size(); q-- > 0;) {
      const int v = order[q];
      const int l = left[(size_t)v], r = right[(size_t)v];
      for (int a = 0; a < 3; ++a) {
        bnd[6 * (size_t)v + a] = mn(bnd[6 * (size_t)l + a], bnd[6 * (size_t)r + a]);
        bnd[6 * (size_t)v + 3 + a] = mx(bnd[6 * (size_t)l + 3 + a], bnd[6 * (size_t)r + 3 + a]);
      }
    }
  }
  auto size_of = [&](int b) { return rlast[(size_t)b] - rfirst[(size_t)b] + 1; };
  auto is_leafish = [&](int b) { return b >= n - 1 || size_of(b) <= S.leaf_max; };
  auto area = [&](int b) {
    const float* q = &bnd[6 * (size_t)b];
    const float dx = q[3] - q[0], dy = q[4] - q[1], dz = q[5] - q[2];
    return dx * dy + dy * dz + dz * dx;
  };
  // breadth-first collapse
  std::vector<int> cur = {0}, depth_cur = {0};
  while (!cur.empty()) {
    std::vector<int> nxt, depth_nxt;
    const size_t base = S.nodes.size();
    const size_t next_base = base + cur.size();
    for (size_t j = 0; j < cur.size(); ++j) {
      const int src = cur[j];
      int ch[WIDTH];
      int c;
      if (is_leafish(src)) {
        ch[0] = src;
        c = 1;
      } else {
        ch[0] = left[(size_t)src];
        ch[1] = right[(size_t)src];
        c = 2;
        while (c < WIDTH) {
          int best = -1;
          float best_a = -1.0f;
          for (int k = 0; k < c; ++k) {
            if (is_leafish(ch[k])) continue;
            const float a = area(ch[k]);
            if (a > best_a) { best_a = a; best = k; }
          }
          if (best < 0) break;
          const int v = ch[best];
          for (int k = c; k > best + 1; --k) ch[k] = ch[k - 1];
          ch[best] = left[(size_t)v];
          ch[best + 1] = right[(size_t)v];
          ++c;
        }
      }
      int k_int = 0;
      for (int k = 0; k < c; ++k) k_int += is_leafish(ch[k]) ? 0 : 1;
      Node node;
      std::memset(&node, 0, sizeof(node));
      for (int s = 0; s < WIDTH; ++s) {
        NodeHalf& hf = node.of(s);
        const int e = s % 4;
        if (s < c) {
          const float* q = &bnd[6 * (size_t)ch[s]];
          hf.lox[e] = q[0]; hf.loy[e] = q[1]; hf.loz[e] = q[2];
          hf.hix[e] = q[3]; hf.hiy[e] = q[4]; hf.hiz[e] = q[5];
          if (!is_leafish(ch[s])) {
            hf.child[e] = (int32_t)(next_base + nxt.size());
            nxt.push_back(ch[s]);
            depth_nxt.push_back(depth_cur[j] + (k_int - 1));
          } else {
            hf.child[e] = (int32_t)(0x80000000u | ((uint32_t)(size_of(ch[s]) - 1) << 27) |
                                    (uint32_t)rfirst[(size_t)ch[s]]);
          }
        } else {
          hf.lox[e] = hf.hix[e] = hf.loy[e] = hf.hiy[e] = hf.loz[e] = hf.hiz[e] = EMPTY_BOX;
          hf.child[e] = EMPTY;
        }
      }
      S.stack_bound = std::max(S.stack_bound, depth_cur[j] + std::max(k_int - 1, 0));
      S.nodes.push_back(node);
    }
    cur.swap(nxt);
    depth_cur.swap(depth_nxt);
    ++S.levels;
  }
  S.stris.resize(9 * (size_t)n);
  for (int k = 0; k < n; ++k)
    std::memcpy(&S.stris[9 * (size_t)k], &S.pos[9 * (size_t)S.ids[(size_t)k]], 9 * sizeof(float));
}

struct HitO {
  float t, b1, b2;
  uint32_t id;
};

void consider(HitO& h, uint32_t id, float t, float b1, float b2) {
  if (t < h.t || (t == h.t && id < h.id)) {
    h.t = t; h.b1 = b1; h.b2 = b2; h.id = id;
  }
}

HitO trace_brute(const Scene& S, const RayO& r) {
  const Pre p = prepare(r);
  HitO h{p.tmax, 0.0f, 0.0f, 0xffffffffu};
  for (int i = 0; i < S.n; ++i) {
    float t, b1, b2;
    if (hit_triangle(p, &S.pos[9 * (size_t)i], &t, &b1, &b2)) consider(h, (uint32_t)i, t, b1, b2);
  }
  return h;
}

// One slab; an axis-parallel ray (|d| < 1e-30) is inside it iff lo <= o <= hi.
void slab_axis(const Pre& r, int a, float lo, float hi, float* t0, float* t1) {
  if (r.par[a]) {
    *t0 = lo <= r.o[a] ? -INFINITY : INFINITY;
    *t1 = hi >= r.o[a] ? INFINITY : -INFINITY;
  } else {
    *t0 = (lo - r.o[a]) * r.inv[a];
    *t1 = (hi - r.o[a]) * r.inv[a];
  }
}

float slab(const Pre& r, const Node& nd, int s, float limit) {
  const NodeHalf& hf = nd.of(s);
  const int e = s % 4;
  float x0, x1, y0, y1, z0, z1;
  slab_axis(r, 0, hf.lox[e], hf.hix[e], &x0, &x1);
  slab_axis(r, 1, hf.loy[e], hf.hiy[e], &y0, &y1);
  slab_axis(r, 2, hf.loz[e], hf.hiz[e], &z0, &z1);
  const float tn = std::fmax(std::fmax(std::fmin(x0, x1), std::fmin(y0, y1)), std::fmax(std::fmin(z0, z1), r.tmin));
  const float tf = std::fmin(std::fmin(std::fmax(x0, x1), std::fmax(y0, y1)), std::fmax(z0, z1)) * 1.00000048f;
  return tn <= std::fmin(tf, limit) ? tn : INFINITY;
}

// Same traversal order as the GPU kernel (leaves inline, internal children
// near-to-far), so node and triangle visit counts match it per ray.
HitO trace_bvh(const Scene& S, const RayO& r, uint64_t* nodes_visited, uint64_t* tris_tested) {
  const Pre p = prepare(r);
  HitO h{p.tmax, 0.0f, 0.0f, 0xffffffffu};
  if (S.nodes.empty()) return h;
  std::vector<int> stack;
  stack.reserve(64);
  int node = 0;
  for (;;) {
    const Node& nd = S.nodes[(size_t)node];
    ++*nodes_visited;
    float te[WIDTH];
    // all entry distances first, with the culling limit of the hit so far:
    // cull only boxes entered beyond it by more than 2^-16 relative
    for (int s = 0; s < WIDTH; ++s) te[s] = slab(p, nd, s, h.t * 1.0000152587890625f);
    int ni = 0, ic[WIDTH];
    float it[WIDTH];
    for (int s = 0; s < WIDTH; ++s) {
      const int32_t c = nd.child(s);
      if (c == EMPTY || !(te[s] < INFINITY)) continue;
      if (c < 0) {
        const uint32_t first = (uint32_t)c & 0x07ffffffu, cnt = (((uint32_t)c >> 27) & 0xfu) + 1u;
        for (uint32_t q = 0; q < cnt; ++q) {
          ++*tris_tested;
          float t, b1, b2;
          if (hit_triangle(p, &S.stris[9 * (size_t)(first + q)], &t, &b1, &b2))
            consider(h, S.ids[first + q], t, b1, b2);
        }
      } else {
        ic[ni] = c;
        it[ni] = te[s];
        ++ni;
      }
    }
    if (ni == 0) {
      if (stack.empty()) break;
      node = stack.back();
      stack.pop_back();
      continue;
    }
    for (int a = 1; a < ni; ++a) {
      const float tv = it[a];
      const int cv = ic[a];
      int b = a - 1;
      while (b >= 0 && it[b] > tv) { it[b + 1] = it[b]; ic[b + 1] = ic[b]; --b; }
      it[b + 1] = tv;
      ic[b + 1] = cv;
    }
    for (int a = ni - 1; a >= 1; --a) stack.push_back(ic[a]);
    node = ic[0];
  }
  return h;
}

F3 sky(int mode, const float d_in[3]) {
  const F3 d = normalize(f3(d_in[0], d_in[1], d_in[2]));
  if (mode == 1) {  // simple_sky, shaders.cu:749-756
    const float mix = std::fmin(std::fmax(0.0f, (o_asin(d.y) * 2.0f) / PI_F), 1.0f);
    const F3 upper = f3(1.0f / 255.0f, 31.0f / 255.0f, 117.0f / 255.0f);
    const F3 lower = scale(f3(143.0f / 255.0f, 179.0f / 255.0f, 203.0f / 255.0f), 0.8f);
    return add(scale(lower, 1.0f - mix), scale(upper, mix));
  }
  // default_background, shaders.cu:740-747
  if (std::fabs(d.x) < 0.01f || std::fabs(d.y) < 0.01f || std::fabs(d.z) < 0.01f) return f3(0, 0, 0);
  return f3((o_atan2(d.z, d.x) + PI_F) / (PI_F * 2.0f), (o_asin(d.y) + PI_F / 2.0f) / PI_F, 0.0f);
}

// Texel coordinate along one axis: repeat wraps modulo n, clamp holds the
// edge texel, mirrored repeat reflects every other period.
int axis_texel(int i, int n, int addr) {
  switch (addr) {
    case 1: return std::min(std::max(i, 0), n - 1);
    case 2: {
      const int r = ((i % (2 * n)) + 2 * n) % (2 * n);
      return r >= n ? 2 * n - 1 - r : r;
    }
    default: return ((i % n) + n) % n;
  }
}

F3 texel(const Scene& S, const int32_t* T, int i, int j) {
  const uint32_t v = S.texels[(size_t)T[5] + (size_t)j * (size_t)T[0] + (size_t)i];
  return f3((float)(v & 255u) / 255.0f, (float)(v >> 8 & 255u) / 255.0f, (float)(v >> 16 & 255u) / 255.0f);
}

// tex2D(normalised coords, normalised float read) for the reference's
// samplers (MulticamScene.cpp:801-834): the coordinate is first reduced to one
// period of its address mode; point filtering takes texel floor(u*W); linear
// filtering weights the 2x2 texels around (u*W - 0.5) by fractions rounded to
// multiples of 1/256.
F3 texture_rgb(const Scene& S, uint32_t t, float u, float v) {
  const int32_t* T = &S.tex[6 * (size_t)t];
  auto reduce = [](float c, int addr) -> float {
    if (std::isnan(c)) return 0.0f;
    if (addr == 1) return std::fmin(std::fmax(c, 0.0f), 1.0f);
    if (addr == 2) return c - 2.0f * std::floor(c * 0.5f);
    return c - std::floor(c);
  };
  const float x = reduce(u, T[2]) * (float)T[0];
  const float y = reduce(v, T[3]) * (float)T[1];
  if (T[4] == 0)
    return texel(S, T, axis_texel((int)std::floor(x), T[0], T[2]), axis_texel((int)std::floor(y), T[1], T[3]));
  const float xs = x - 0.5f, ys = y - 0.5f;
  const float x0 = std::floor(xs), y0 = std::floor(ys);
  const float wx = std::nearbyint((xs - x0) * 256.0f) * (1.0f / 256.0f);
  const float wy = std::nearbyint((ys - y0) * 256.0f) * (1.0f / 256.0f);
  const int ix = (int)x0, iy = (int)y0;
  const int xa = axis_texel(ix, T[0], T[2]), xb = axis_texel(ix + 1, T[0], T[2]);
  const int ya = axis_texel(iy, T[1], T[3]), yb = axis_texel(iy + 1, T[1], T[3]);
  const F3 r0 = add(scale(texel(S, T, xa, ya), 1.0f - wx), scale(texel(S, T, xb, ya), wx));
  const F3 r1 = add(scale(texel(S, T, xa, yb), 1.0f - wx), scale(texel(S, T, xb, yb), wx));
  return add(scale(r0, 1.0f - wy), scale(r1, wy));
}

F3 shade(const Scene& S, const HitO& h, const RayO& r) {
  if (h.id == 0xffffffffu) return sky(S.sky, r.d);
  const float* c = &S.col[9 * (size_t)h.id];
  const uint32_t m = S.mode[h.id];
  if ((m & 3u) == 1u) return f3(c[0], c[1], c[2]);  // material base colour, no linearize
  if ((m & 3u) == 2u) {
    // textured: UV interpolated as in LocalGeometry.h:92-103, then linearize (shaders.cu:801-805)
    const float w0 = (1.0f - h.b1) - h.b2;
    const float u = (w0 * c[0] + h.b1 * c[2]) + h.b2 * c[4];
    const float v = (w0 * c[1] + h.b1 * c[3]) + h.b2 * c[5];
    const F3 t = texture_rgb(S, m >> 2, u, v);
    return f3(o_pow(t.x, 2.2f), o_pow(t.y, 2.2f), o_pow(t.z, 2.2f));
  }
  const float w0 = (1.0f - h.b1) - h.b2;                 // LocalGeometry.h:148
  F3 v;
  v.x = (w0 * c[0] + h.b1 * c[3]) + h.b2 * c[6];
  v.y = (w0 * c[1] + h.b1 * c[4]) + h.b2 * c[7];
  v.z = (w0 * c[2] + h.b1 * c[5]) + h.b2 * c[8];
  return f3(o_pow(v.x, 2.2f), o_pow(v.y, 2.2f), o_pow(v.z, 2.2f));   // linearize, shaders.cu:93-100, :796
}

template <typename F>
void parallel_for(size_t n, int threads, F fn) {
  if (threads <= 1 || n < 64) {
    for (size_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<size_t> next(0);
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        const size_t b = next.fetch_add(16);
        if (b >= n) break;
        const size_t e = std::min(n, b + 16);
        for (size_t i = b; i < e; ++i) fn(i);
      }
    });
  for (auto& th : pool) th.join();
}

PoseO pose_from(const float* p12) {
  PoseO P;
  std::memcpy(&P, p12, sizeof(PoseO));
  return P;
}

uint32_t samples_of(const uint32_t* counts, int spp, size_t i) { return counts ? counts[i] : (uint32_t)spp; }

}  // namespace

extern "C" {

void oracle_philox(const uint32_t* ctr, uint32_t k0, uint32_t k1, uint32_t* out) { philox(ctr, k0, k1, out); }

// kind: 0 sin, 1 cos, 2 log, 3 exp2, 4 pow(x, y), 5 asin, 6 atan2(x, y)
void oracle_math(int kind, const float* x, const float* y, size_t n, float* out) {
  for (size_t i = 0; i < n; ++i) {
    float s, c;
    switch (kind) {
      case 0: o_sincos(x[i], &s, &c); out[i] = s; break;
      case 1: o_sincos(x[i], &s, &c); out[i] = c; break;
      case 2: out[i] = o_log(x[i]); break;
      case 3: out[i] = o_exp2(x[i]); break;
      case 4: out[i] = o_pow(x[i], y[i]); break;
      case 5: out[i] = o_asin(x[i]); break;
      default: out[i] = o_atan2(x[i], y[i]); break;
    }
  }
}

void* oracle_scene_create(const float* pos, const float* col, const uint32_t* mode, int64_t n, int sky,
                          int leaf_max) {
  Scene* S = new Scene();
  S->n = (int)n;
  S->sky = sky;
  S->leaf_max = leaf_max;
  S->pos.assign(pos, pos + 9 * (size_t)n);
  S->col.assign(col, col + 9 * (size_t)n);
  S->mode.assign(mode, mode + (size_t)n);
  build(*S);
  return S;
}

void oracle_scene_free(void* h) { delete static_cast<Scene*>(h); }

// Base-colour textures of a scene: 6 ints per texture (width, height, wrap s,
// wrap t, filter, first texel) and the RGBA8 texels.
void oracle_scene_set_textures(void* h, const int32_t* desc, int64_t ntex, const uint32_t* texels, int64_t ntexels) {
  Scene* S = static_cast<Scene*>(h);
  S->tex.assign(desc, desc + 6 * (size_t)ntex);
  S->texels.assign(texels, texels + (size_t)ntexels);
}

// One texture lookup (texture t at (u, v)) for direct checks of the sampler.
void oracle_texture_sample(void* h, uint32_t t, float u, float v, float* rgb) {
  const F3 c = texture_rgb(*static_cast<Scene*>(h), t, u, v);
  rgb[0] = c.x;
  rgb[1] = c.y;
  rgb[2] = c.z;
}

// info: tris, nodes, stack bound, levels, leaf size, centroid lo xyz/hi xyz as float bits (6),
// node bytes
void oracle_scene_info(void* h, int64_t* info) {
  const Scene& S = *static_cast<Scene*>(h);
  info[0] = S.n;
  info[1] = (int64_t)S.nodes.size();
  info[2] = S.stack_bound;
  info[3] = S.levels;
  info[4] = S.leaf_max;
  info[11] = (int64_t)sizeof(Node);
  for (int a = 0; a < 3; ++a) {
    info[5 + a] = S.n ? (int64_t)bits(S.clo[a]) : 0;
    info[8 + a] = S.n ? (int64_t)bits(S.chi[a]) : 0;
  }
}

void oracle_scene_nodes(void* h, void* dst) {
  const Scene& S = *static_cast<Scene*>(h);
  if (!S.nodes.empty()) std::memcpy(dst, S.nodes.data(), S.nodes.size() * sizeof(Node));
}

void oracle_scene_order(void* h, uint32_t* ids, uint32_t* codes) {
  const Scene& S = *static_cast<Scene*>(h);
  if (S.n == 0) return;
  std::memcpy(ids, S.ids.data(), 4 * (size_t)S.n);
  std::memcpy(codes, S.codes.data(), 4 * (size_t)S.n);
}

// rays: 8 floats each.  Ray k = (pose p, omm i, sample s) in (pose, omm, sample) order.
void oracle_generate_rays(const float* omms, int64_t n_omm, const uint32_t* counts, int spp, const float* poses,
                          int64_t n_pose, uint32_t seed, uint32_t eye, uint32_t frame, float* rays) {
  uint64_t k = 0;
  for (int64_t p = 0; p < n_pose; ++p) {
    const PoseO P = pose_from(poses + 12 * p);
    for (int64_t i = 0; i < n_omm; ++i) {
      Omm om;
      std::memcpy(&om, omms + 8 * i, sizeof(Omm));
      const uint32_t n = samples_of(counts, spp, (size_t)i);
      for (uint32_t s = 0; s < n; ++s, ++k) {
        const RayO r = make_ray(om, P, seed, eye, (uint32_t)i, s, frame, (uint32_t)p);
        std::memcpy(rays + 8 * k, &r, sizeof(RayO));
      }
    }
  }
}

// mode 0: brute force, 1: BVH.  ids: triangle id or -1; tbb: (t, b1, b2).
// visits (nullable): per ray (nodes visited, triangles tested).
void oracle_trace(void* h, const float* rays, int64_t n, int mode, int32_t* ids, float* tbb, uint64_t* visits,
                  int threads) {
  const Scene& S = *static_cast<Scene*>(h);
  parallel_for((size_t)n, threads, [&](size_t k) {
    RayO r;
    std::memcpy(&r, rays + 8 * k, sizeof(RayO));
    uint64_t nv = 0, nt = 0;
    const HitO hh = mode == 0 ? trace_brute(S, r) : trace_bvh(S, r, &nv, &nt);
    ids[k] = hh.id == 0xffffffffu ? -1 : (int32_t)hh.id;
    tbb[3 * k] = hh.t;
    tbb[3 * k + 1] = hh.b1;
    tbb[3 * k + 2] = hh.b2;
    if (visits) {
      visits[2 * k] = nv;
      visits[2 * k + 1] = nt;
    }
  });
}

// Renders one eye at n_pose poses.  Outputs (each nullable):
//   samples  4 floats per ray (r, g, b, t or -1), (pose, omm, sample) order
//   col_gpu  3 floats per (pose, omm): canonical wave order
//   col_seq  3 floats per (pose, omm): the reference's sequential order
//   col_f64  3 doubles per (pose, omm): FP64 accumulation of unscaled samples / n
//   visits   2 uint64 totals (nodes visited, triangles tested)
//   counts_per_unit: counts holds one entry per (pose, omm) instead of per omm
//   sample_base:     sample index of each unit's first sample (top-up passes)
void oracle_render(void* h, const float* omms, int64_t n_omm, const uint32_t* counts, int counts_per_unit, int spp,
                   const float* poses, int64_t n_pose, uint32_t seed, uint32_t eye, uint32_t frame,
                   uint32_t sample_base, float* samples, float* col_gpu, float* col_seq, double* col_f64,
                   uint64_t* visits, int threads) {
  const Scene& S = *static_cast<Scene*>(h);
  const size_t units = (size_t)(n_omm * n_pose);
  auto count_of = [&](size_t u) {
    return counts_per_unit ? samples_of(counts, spp, u) : samples_of(counts, spp, u % (size_t)n_omm);
  };
  std::vector<uint64_t> first(units + 1, 0);
  for (size_t u = 0; u < units; ++u) first[u + 1] = first[u] + count_of(u);
  std::atomic<uint64_t> tot_nodes(0), tot_tris(0);
  parallel_for(units, threads, [&](size_t u) {
    const size_t p = u / (size_t)n_omm, i = u % (size_t)n_omm;
    const PoseO P = pose_from(poses + 12 * p);
    Omm om;
    std::memcpy(&om, omms + 8 * i, sizeof(Omm));
    const uint32_t n = count_of(u);
    const float inv = 1.0f / (float)(n > 0 ? n : 1u);
    float lane[64][3] = {};
    float seq[3] = {0.0f, 0.0f, 0.0f};
    double f64[3] = {0.0, 0.0, 0.0};
    uint64_t nv = 0, nt = 0;
    for (uint32_t s = 0; s < n; ++s) {
      const RayO r = make_ray(om, P, seed, eye, (uint32_t)i, sample_base + s, frame, (uint32_t)p);
      const HitO hh = trace_bvh(S, r, &nv, &nt);
      const F3 c = shade(S, hh, r);
      if (samples) {
        float* o = samples + 4 * (first[u] + s);
        o[0] = c.x; o[1] = c.y; o[2] = c.z;
        o[3] = hh.id == 0xffffffffu ? -1.0f : hh.t;
      }
      const float sc[3] = {c.x * inv, c.y * inv, c.z * inv};   // shaders.cu:730
      for (int a = 0; a < 3; ++a) {
        lane[s & 63][a] = lane[s & 63][a] + sc[a];
        seq[a] = seq[a] + sc[a];                               // shaders.cu:341-347
      }
      f64[0] += c.x; f64[1] += c.y; f64[2] += c.z;
    }
    for (int m = 32; m >= 1; m >>= 1)
      for (int l = 0; l < 64; ++l) {
        if (l & m) continue;
        for (int a = 0; a < 3; ++a) {
          const float v = lane[l][a] + lane[l ^ m][a];
          lane[l][a] = v;
          lane[l ^ m][a] = v;
        }
      }
    if (col_gpu) for (int a = 0; a < 3; ++a) col_gpu[3 * u + a] = lane[0][a];
    if (col_seq) for (int a = 0; a < 3; ++a) col_seq[3 * u + a] = seq[a];
    if (col_f64) for (int a = 0; a < 3; ++a) col_f64[3 * u + a] = n ? f64[a] / (double)n : 0.0;
    tot_nodes += nv;
    tot_tris += nt;
  });
  if (visits) {
    visits[0] = tot_nodes.load();
    visits[1] = tot_tris.load();
  }
}

// setCameraPose, libEyeRenderer.cpp:380-388: reset, rotate about world X, Y,
// Z (rotatePoint with a normalised axis, DataRecordCamera.h:66-87), translate.
// out: 12 floats (position, X axis, Y axis, Z axis).
void oracle_pose(const float* pos, const float* rot, float* out) {
  F3 ax[3] = {f3(1, 0, 0), f3(0, 1, 0), f3(0, 0, 1)};
  const F3 world[3] = {f3(1, 0, 0), f3(0, 1, 0), f3(0, 0, 1)};
  for (int r = 0; r < 3; ++r) {
    const F3 k = normalize(world[r]);
    float s, c;
    o_sincos(rot[r], &s, &c);
    for (int a = 0; a < 3; ++a)
      ax[a] = add(add(scale(ax[a], c), scale(cross(k, ax[a]), s)), scale(k, (1.0f - c) * dot(k, ax[a])));
  }
  out[0] = 0.0f + pos[0]; out[1] = 0.0f + pos[1]; out[2] = 0.0f + pos[2];
  for (int a = 0; a < 3; ++a) { out[3 + 3 * a] = ax[a].x; out[4 + 3 * a] = ax[a].y; out[5 + 3 * a] = ax[a].z; }
}

// ------------------------------------------------------ learned sampler ----
// FP64 16-64-64-64-1 ReLU MLP with the GPU kernels' parameter order
// (W1[16][64], b1, W2[64][64], b2, W3[64][64], b3, W4[64], b4; W[in][out]).
// mode 0: exact FP64.  mode 1: BF16 rounding (nearest even) of the inputs,
// weights, hidden activations and pre-activation gradients where the GPU
// feeds them to MFMA, FP64 sums elsewhere.
static double bf16r(double x) {
  const float f = (float)x;
  uint32_t u = bits(f);
  u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
  return (double)from_bits(u);
}

void oracle_mlp_loss_grad(const float* params, const float* X, const float* T, int64_t n, int mode, double* y_out,
                          double* grad, double* loss_out) {
  const int H = 64, IN = 16;
  auto R = [&](double v) { return mode ? bf16r(v) : v; };
  const float* W1 = params;
  const float* b1 = W1 + IN * H;
  const float* W2 = b1 + H;
  const float* b2 = W2 + H * H;
  const float* W3 = b2 + H;
  const float* b3 = W3 + H * H;
  const float* W4 = b3 + H;
  const float b4 = W4[H];
  std::vector<double> w1(IN * H), w2(H * H), w3(H * H), w4(H);
  for (int k = 0; k < IN * H; ++k) w1[k] = R(W1[k]);
  for (int k = 0; k < H * H; ++k) { w2[k] = R(W2[k]); w3[k] = R(W3[k]); }
  for (int k = 0; k < H; ++k) w4[k] = R(W4[k]);
  const int P = IN * H + H + H * H + H + H * H + H + H + 1;
  if (grad) std::fill(grad, grad + P, 0.0);
  double* gW1 = grad;
  double* gb1 = grad ? gW1 + IN * H : nullptr;
  double* gW2 = grad ? gb1 + H : nullptr;
  double* gb2 = grad ? gW2 + H * H : nullptr;
  double* gW3 = grad ? gb2 + H : nullptr;
  double* gb3 = grad ? gW3 + H * H : nullptr;
  double* gW4 = grad ? gb3 + H : nullptr;
  double loss = 0.0;
  std::vector<double> x(IN), h1(H), h2(H), h3(H), d3(H), d2(H), d1(H);
  for (int64_t s = 0; s < n; ++s) {
    for (int i = 0; i < IN; ++i) x[i] = R(X[s * IN + i]);
    auto layer = [&](const std::vector<double>& in, int nin, const std::vector<double>& w, const float* b,
                     std::vector<double>& out) {
      for (int o = 0; o < H; ++o) {
        double z = 0.0;
        for (int i = 0; i < nin; ++i) z += in[i] * w[i * H + o];
        z = mode ? (double)((float)z + b[o]) : z + b[o];
        out[o] = R(z > 0.0 ? z : 0.0);
      }
    };
    layer(x, IN, w1, b1, h1);
    layer(h1, H, w2, b2, h2);
    layer(h2, H, w3, b3, h3);
    double y = b4;
    for (int f = 0; f < H; ++f) y += h3[f] * w4[f];
    if (y_out) y_out[s] = y;
    if (!T) continue;
    const double err = y - T[s];
    loss += err * err;
    if (!grad) continue;
    const double dy = 2.0 * err / (double)n;
    gW4[H] += dy;
    for (int f = 0; f < H; ++f) {
      gW4[f] += dy * h3[f];
      d3[f] = h3[f] > 0.0 ? w4[f] * dy : 0.0;
      gb3[f] += d3[f];
    }
    auto back = [&](const std::vector<double>& d, const std::vector<double>& hin, int nin,
                    const std::vector<double>& w, double* gW, std::vector<double>* dprev, double* gbprev) {
      for (int i = 0; i < nin; ++i)
        for (int o = 0; o < H; ++o) gW[i * H + o] += hin[i] * R(d[o]);
      if (!dprev) return;
      for (int i = 0; i < nin; ++i) {
        double acc = 0.0;
        for (int o = 0; o < H; ++o) acc += w[i * H + o] * R(d[o]);
        (*dprev)[i] = hin[i] > 0.0 ? acc : 0.0;
        gbprev[i] += (*dprev)[i];
      }
    };
    back(d3, h2, H, w3, gW3, &d2, gb2);
    back(d2, h1, H, w2, gW2, &d1, gb1);
    back(d1, x, IN, w1, gW1, nullptr, nullptr);
  }
  if (loss_out) *loss_out = T ? loss / (double)n : 0.0;
}

// make_color, shaders.cu:180-189: 4 bytes per colour.
void oracle_make_color(const float* rgb, int64_t n, uint8_t* out) {
  const float g = (float)(1.0 / (double)2.2f);
  for (int64_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      const float c = std::fmin(std::fmax(rgb[3 * i + a], 0.0f), 1.0f);
      out[4 * i + a] = (uint8_t)(o_pow(c, g) * 255.0f);
    }
    out[4 * i + 3] = 255;
  }
}

// Primary rays of the regular cameras, libEyeRenderer3/shaders.cu:198-333.
// pose: position, x, y, z axes (12 floats); kind 0 pinhole (__raygen__pinhole:
// direction z*sz + d.x*x*sx + d.y*y*sy, unnormalised), 1 orthographic
// (direction z, origin shifted by d.x*x*sx + d.y*y*sy), 2 panoramic (unit
// sphere direction in the camera axes, origin moved out by start_radius).
void oracle_camera_rays(int kind, const float* pose, const float* sc, float start_radius, int w, int h,
                        float* out) {
  const F3 O = f3(pose[0], pose[1], pose[2]), X = f3(pose[3], pose[4], pose[5]);
  const F3 Y = f3(pose[6], pose[7], pose[8]), Z = f3(pose[9], pose[10], pose[11]);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const float fx = (float)x / (float)w, fy = (float)y / (float)h;
      const float d0 = 2.0f * fx - 1.0f, d1 = 2.0f * fy - 1.0f;
      F3 o, d;
      if (kind == 1) {
        d = Z;
        o = add(add(O, scale(scale(X, d0), sc[0])), scale(scale(Y, d1), sc[1]));
      } else if (kind == 2) {
        const float a0 = d0 * -PI_F + PI_F / 2.0f, a1 = d1 * (PI_F / 2.0f) + 0.0f;
        float s0, c0, s1, c1;
        o_sincos(a0, &s0, &c0);
        o_sincos(a1, &s1, &c1);
        const F3 u = f3(c0 * c1, s1, s0 * c1);
        d = normalize(add(add(scale(X, u.x), scale(Y, u.y)), scale(Z, u.z)));
        o = add(O, scale(d, start_radius));
      } else {
        d = add(add(scale(Z, sc[2]), scale(scale(X, d0), sc[0])), scale(scale(Y, d1), sc[1]));
        o = O;
      }
      float* r = out + 8 * ((size_t)y * (size_t)w + (size_t)x);
      r[0] = o.x; r[1] = o.y; r[2] = o.z; r[3] = 0.01f;
      r[4] = d.x; r[5] = d.y; r[6] = d.z; r[7] = 1e16f;
    }
}

// Radiance of arbitrary rays: closest hit (BVH) shaded, or the sky.
void oracle_radiance(void* h, const float* rays, int64_t n, float* rgb, int threads) {
  const Scene& S = *static_cast<Scene*>(h);
  parallel_for((size_t)n, threads, [&](size_t k) {
    RayO r;
    std::memcpy(&r, rays + 8 * k, sizeof(RayO));
    uint64_t nv = 0, nt = 0;
    const HitO hh = trace_bvh(S, r, &nv, &nt);
    const F3 c = shade(S, hh, r);
    rgb[3 * k] = c.x;
    rgb[3 * k + 1] = c.y;
    rgb[3 * k + 2] = c.z;
  });
}

// Spherical projections, libEyeRenderer3/shaders.cu:412-640: for each pixel
// of a w x h window, the index of the ommatidium whose position (variant 0)
// or direction (1; 2 = split by the sign of position.x against the pixel's
// half) makes the smallest angle with the pixel's sphere point.  The angle
// order is taken on the cosine (acos is decreasing), first index on ties.
void oracle_spherical_ids(const float* omm, int64_t n, int w, int h, int variant, uint32_t* ids) {
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const float fx = (float)x / (float)w, fy = (float)y / (float)h;
      float d0 = 2.0f * fx - 1.0f, d1 = 2.0f * fy - 1.0f;
      if (variant == 2) {
        float sx = fx * 2.0f;
        if (sx > 1.0f) sx = sx - 1.0f;
        else sx = sx - 0.0f;
        d0 = sx * 2.0f - 1.0f;
        d1 = fy * 2.0f - 1.0f;
      }
      const float a0 = d0 * -PI_F + PI_F / 2.0f;
      const float a1 = d1 * (PI_F / 2.0f) + 0.0f;
      float s0, c0, s1, c1;
      o_sincos(a0, &s0, &c0);
      o_sincos(a1, &s1, &c1);
      const F3 sp = f3(c0 * c1, s1, s0 * c1);
      const float ls = std::sqrt(dot(sp, sp));
      uint32_t best_i = 0;
      float best = 0.0f;
      for (int64_t i = 0; i < n; ++i) {
        const float* o = omm + 8 * i;
        const F3 v = variant == 0 ? f3(o[0], o[1], o[2]) : f3(o[3], o[4], o[5]);
        const float c = dot(v, sp) / (std::sqrt(dot(v, v)) * ls);
        if (i == 0) {
          best = c;
          continue;
        }
        if (variant == 2 && !((o[0] > 0.0f && fx > 0.5f) || (o[0] < 0.0f && fx < 0.5f))) continue;
        if (c > best) {
          best = c;
          best_i = (uint32_t)i;
        }
      }
      ids[(size_t)y * (size_t)w + (size_t)x] = best_i;
    }
}

}  // extern "C"
