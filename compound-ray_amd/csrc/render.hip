// gfx950 render kernels: fused Gaussian-cone ray generation + wide BVH
// traversal + shading (replacing __raygen__ommatidium, the OptiX traversal,
// __closesthit__radiance and the miss programs, libEyeRenderer3/shaders.cu:664-811),
// and the per-ommatidium wave reduction + projection (shaders.cu:341-406).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "render.h"

namespace eye {
namespace {

constexpr int kTraceThreads = 256;
// XCDs of the device: workgroup b runs on XCD b % kXcds, each with its own L2.
constexpr uint32_t kXcds = 8;
static_assert(kQueueWords >= (int)kXcds, "one work counter per XCD slice");
constexpr int kPacketStackMax = 128;  // two VGPRs of 64 lanes

struct SceneView {
  const WideNode* nodes;
  const float4* tris;
  const float4* cols;
  const uint32_t* texels;
  const TexView* textures;
  int node_count;
  int sky_mode;
  uint32_t* fault;   // host-visible word set when a packet traversal is entered by a partial wave
};

SceneView view_of(const DeviceScene& s) {
  SceneView v;
  v.nodes = s.nodes;
  v.tris = s.tris;
  v.cols = s.cols;
  v.texels = s.texels;
  v.textures = s.textures;
  v.node_count = s.node_count;
  v.sky_mode = s.sky_mode;
  v.fault = s.fault;
  return v;
}

struct Hit {
  float t, b1, b2;
  uint32_t k;   // sorted triangle position
  uint32_t id;  // original triangle id, 0xffffffff for a miss
  float det;    // during traversal b1, b2 hold V, W of the closest hit and det its
                // determinant; finish_hit divides them (b1 = V / det, b2 = W / det)
};

__device__ __forceinline__ void finish_hit(Hit& h) {
  if (h.id != 0xffffffffu) {
    h.b1 = h.b1 / h.det;
    h.b2 = h.b2 / h.det;
  }
}

// Pointers into the constant address space: wave-uniform loads through them
// become scalar (SMEM) loads, fetched once per wave into SGPRs.  Only used for
// data no kernel writes while it is read (nodes, triangles).
typedef float nf4 __attribute__((ext_vector_type(4)));
typedef int ni4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(4))) const nf4 kf4;
typedef __attribute__((address_space(4))) const ni4 ki4;

struct UniformPtr {   // float4 array read through the constant address space
  const void* p;
};
struct VaryingPtr {   // float4 array read with per-lane (vector) loads
  const float4* p;
};
__device__ __forceinline__ float4 load4(UniformPtr b, size_t i) {
  const nf4 v = ((const kf4*)b.p)[i];
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 load4(VaryingPtr b, size_t i) { return b.p[i]; }
__device__ __forceinline__ int4 load_int4_uniform(const void* p, size_t i) {
  const ni4 v = ((const ki4*)p)[i];
  return make_int4(v.x, v.y, v.z, v.w);
}

template <typename P4>
__device__ __forceinline__ void test_leaf(P4 tris, int32_t code, const RayPre& rp, Hit& h) {
  const uint32_t first = leaf_first(code);
  const uint32_t cnt = leaf_count(code);
  for (uint32_t q = 0; q < cnt; ++q) {
    const uint32_t k = first + q;
    const float4 a = load4(tris, 3 * (size_t)k + 0);
    const float4 b = load4(tris, 3 * (size_t)k + 1);
    const float4 c = load4(tris, 3 * (size_t)k + 2);
    float t, vv, ww, det;
    if (intersect_triangle_raw(rp, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), &t, &vv, &ww, &det)) {
      const uint32_t id = __float_as_uint(a.w);
      if (t < h.t || (t == h.t && id < h.id)) {
        h.t = t;
        h.b1 = vv;
        h.b2 = ww;
        h.det = det;
        h.k = k;
        h.id = id;
      }
    }
  }
}

__device__ __forceinline__ float lane_f(int v, int l) { return __int_as_float(__builtin_amdgcn_readlane(v, l)); }

// A wave-uniform leaf (packet traversal, every lane of the wave executing).
// Triangle records are 12 dwords (three float4: vertex xyz, and the original
// id in the first record's w).  One coalesced load puts dword j of up to five
// consecutive records on lane j; v_readlane then hands each record to the
// scalar unit.  A leaf of n triangles costs ceil(n / 5) dependent memory round
// trips instead of n, and the tested values are the same bits.
__device__ __forceinline__ void test_leaf_wave(const float4* tris, int32_t code, const RayPre& rp, Hit& h) {
  constexpr uint32_t kPerLoad = 5;
  const uint32_t first = leaf_first(code);
  const uint32_t cnt = leaf_count(code);
  const int lane = (int)(threadIdx.x & 63);
  const int* words = reinterpret_cast<const int*>(tris + 3 * (size_t)first);
  for (uint32_t q0 = 0; q0 < cnt; q0 += kPerLoad) {
    const uint32_t m = cnt - q0 < kPerLoad ? cnt - q0 : kPerLoad;
    const int v = lane < (int)(12 * m) ? words[12 * (size_t)q0 + lane] : 0;
    for (uint32_t q = 0; q < m; ++q) {
      const int o = (int)(12 * q);
      const V3 a = v3(__int_as_float(__builtin_amdgcn_readlane(v, o + 0)),
                      __int_as_float(__builtin_amdgcn_readlane(v, o + 1)),
                      __int_as_float(__builtin_amdgcn_readlane(v, o + 2)));
      const uint32_t id = (uint32_t)__builtin_amdgcn_readlane(v, o + 3);
      const V3 b = v3(__int_as_float(__builtin_amdgcn_readlane(v, o + 4)),
                      __int_as_float(__builtin_amdgcn_readlane(v, o + 5)),
                      __int_as_float(__builtin_amdgcn_readlane(v, o + 6)));
      const V3 c = v3(__int_as_float(__builtin_amdgcn_readlane(v, o + 8)),
                      __int_as_float(__builtin_amdgcn_readlane(v, o + 9)),
                      __int_as_float(__builtin_amdgcn_readlane(v, o + 10)));
      float t, vv, ww, det;
      if (intersect_triangle_raw(rp, a, b, c, &t, &vv, &ww, &det)) {
        if (t < h.t || (t == h.t && id < h.id)) {
          h.t = t;
          h.b1 = vv;
          h.b2 = ww;
          h.det = det;
          h.k = first + q0 + q;
          h.id = id;
        }
      }
    }
  }
}

// Inserts (key, child) into the list of the first n entries of (keys,
// children), kept in increasing key order; among equal keys the earlier
// insertion stays first.  Every index is a compile-time constant, so the
// lists stay in registers (scalar registers when the values are
// wave-uniform): a list indexed by a run-time count would be placed in scratch.
template <typename K>
__device__ __forceinline__ void sorted_insert(K (&keys)[kBvhWidth], int (&children)[kBvhWidth], int n, K key,
                                              int child) {
  int pos = 0;
#pragma unroll
  for (int i = 0; i < kBvhWidth; ++i) pos += (i < n && !(key < keys[i])) ? 1 : 0;
#pragma unroll
  for (int j = kBvhWidth - 1; j >= 0; --j) {
    const K below_k = j > 0 ? keys[j - 1] : key;
    const int below_c = j > 0 ? children[j - 1] : child;
    keys[j] = j < pos ? keys[j] : (j == pos ? key : below_k);
    children[j] = j < pos ? children[j] : (j == pos ? child : below_c);
  }
}

// Closest hit over the kBvhWidth-wide BVH.  `stack` points at this lane's column of
// an LDS array laid out [depth][blockDim.x].  The result is independent of
// the traversal order: the (t, id) pair is minimised lexicographically.  At a
// node, the entry distances of all child boxes use the culling limit of
// the hit found before the node; the slots are taken in order: leaves are
// tested at once, internal children are visited nearest first (equal
// distances in slot order).
template <bool kStats, bool kParallel>
__device__ __forceinline__ void closest_hit_impl(const SceneView& S, const RayPre& rp, uint32_t* stack,
                                                 int stride, Hit& h, uint32_t* nodes_visited,
                                                 uint32_t* tris_tested) {
  h.t = rp.tmax;
  h.b1 = h.b2 = 0.0f;
  h.k = 0xffffffffu;
  h.id = 0xffffffffu;
  if (S.node_count == 0) return;
  int sp = 0;
  int node = 0;
  for (;;) {
    if (kStats) ++*nodes_visited;
    const float4* np = reinterpret_cast<const float4*>(S.nodes + node);
    const float limit = h.t * kCullSlack;
    int ni = 0;
    int ic[kBvhWidth];
    float it[kBvhWidth];
#pragma unroll
    for (int hf = 0; hf < kBvhHalves; ++hf) {
      const float4* q = np + 8 * hf;
      const float4 lx = q[0], hx = q[1], ly = q[2], hy = q[3], lz = q[4], hz = q[5];
      const int4 ch = reinterpret_cast<const int4*>(q)[6];
      float te[kHalfWidth];
      te[0] = slab_enter<kParallel>(rp, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, limit);
      te[1] = slab_enter<kParallel>(rp, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, limit);
      te[2] = slab_enter<kParallel>(rp, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, limit);
      te[3] = slab_enter<kParallel>(rp, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, limit);
      const int32_t cc[kHalfWidth] = {ch.x, ch.y, ch.z, ch.w};
#pragma unroll
      for (int s = 0; s < kHalfWidth; ++s) {
        const int32_t c = cc[s];
        if (c == kChildEmpty || !(te[s] < INFINITY)) continue;
        if (child_is_leaf(c)) {
          if (kStats) *tris_tested += leaf_count(c);
          test_leaf(VaryingPtr{S.tris}, c, rp, h);
        } else {
          sorted_insert(it, ic, ni, te[s], c);
          ++ni;
        }
      }
    }
    if (ni == 0) {
      if (sp == 0) break;
      --sp;
      node = (int)stack[sp * stride];
      continue;
    }
    // internal children are in near-to-far order: push all but the nearest,
    // farthest first, and visit the nearest
#pragma unroll
    for (int a = kBvhWidth - 1; a >= 1; --a) {
      if (a < ni) {
        stack[sp * stride] = (uint32_t)ic[a];
        ++sp;
      }
    }
    node = ic[0];
  }
}

// Rays with an exactly axis-parallel direction take the variant with the
// parallel-slab rule; all others the plain slab test.
template <bool kStats = false>
__device__ __forceinline__ void closest_hit(const SceneView& S, const RayPre& rp, uint32_t* stack, int stride,
                                            Hit& h, uint32_t* nodes_visited = nullptr,
                                            uint32_t* tris_tested = nullptr) {
  if (rp.par != 0)
    closest_hit_impl<kStats, true>(S, rp, stack, stride, h, nodes_visited, tris_tested);
  else
    closest_hit_impl<kStats, false>(S, rp, stack, stride, h, nodes_visited, tris_tested);
  finish_hit(h);
}

__device__ __forceinline__ V3 shade(const SceneView& S, const Hit& h, float dx, float dy, float dz) {
  if (h.id == 0xffffffffu) return sky_colour(S.sky_mode, dx, dy, dz);
  const float4 c0 = S.cols[3 * (size_t)h.k + 0];
  const float4 c1 = S.cols[3 * (size_t)h.k + 1];
  const uint32_t word = __float_as_uint(c0.w);
  if (colour_mode(word) == kTextureColour) {
    const float uv[6] = {c0.x, c0.y, c0.z, c1.x, c1.y, c1.z};
    return shade_texture(h.b1, h.b2, uv, S.texels, S.textures[texture_index(word)]);
  }
  const float4 c2 = S.cols[3 * (size_t)h.k + 2];
  return shade_hit(h.b1, h.b2, v3(c0.x, c0.y, c0.z), v3(c1.x, c1.y, c1.z), v3(c2.x, c2.y, c2.z), word);
}

// Maps a ray number to (pose, ommatidium, sample, samples of that ommatidium).
__device__ __forceinline__ void locate_ray(const EyeLaunch& L, uint64_t r, int* pose, int* omm, int* sample,
                                           int* n) {
  const uint64_t p = r / L.rays_per_pose;
  const uint32_t local = (uint32_t)(r - p * L.rays_per_pose);
  *pose = (int)p;
  if (L.offsets == nullptr) {
    const int i = (int)(local / (uint32_t)L.spp);
    *omm = i;
    *sample = (int)(local - (uint32_t)i * (uint32_t)L.spp);
    *n = L.spp;
    return;
  }
  // last i with offsets[i] <= local (empty ommatidia are skipped naturally)
  int lo = 0, hi = L.omm_count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L.offsets[mid] <= local) lo = mid;
    else hi = mid - 1;
  }
  *omm = lo;
  *sample = (int)(local - L.offsets[lo]);
  *n = (int)(L.offsets[lo + 1] - L.offsets[lo]);
}

// ---------------------------------------------------------------------------
// Wave-packet traversal.  The 64 rays of a wave are samples of one or two
// neighbouring ommatidia: nearly the same origin and direction.  The wave
// walks the BVH together: the node index is wave-uniform (scalar loads), each
// lane tests the children against its own ray and closest hit, a child is
// entered when any lane's ray enters it, and every active lane tests the
// triangles of an entered leaf.  Each lane still computes the exact
// lexicographic (t, id) minimum over the triangles its ray can hit (a lane
// only ever visits a superset of the nodes its own traversal would), so
// results are identical to per-ray traversal whatever the visiting order.
//
// A step is one long chain of dependent instructions, so the step keeps off
// LDS: the wave's stack lives in two VGPRs, entry k in lane k & 63 of
// stack[k >> 6] (a lane select to push, v_readlane to pop; kPacketStackMax
// entries), and entered children are ordered nearest first by the entry
// distance of the first lane that enters them, read with v_readlane and
// compared as bits on the scalar unit (entry distances are >= tmin >= 0,
// where the float order is the integer order).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void stack_push(int& s0, int& s1, int lane, int sp, int v) {
  if (sp < 64)
    s0 = lane == sp ? v : s0;
  else
    s1 = lane == sp - 64 ? v : s1;
}
__device__ __forceinline__ int stack_top(int s0, int s1, int sp) {
  return sp < 64 ? __builtin_amdgcn_readlane(s0, sp) : __builtin_amdgcn_readlane(s1, sp - 64);
}

// One node of a packet step: tests its leaf children (every lane, against
// its own ray) and appends the internal children that any lane enters to
// ic / ik (child, ordering key).
template <bool kParallel, bool kScalar>
__device__ __forceinline__ void packet_visit(const SceneView& S, const RayPre& rp, bool active, int node, Hit& h,
                                             int (&ic)[kBvhWidth], uint32_t (&ik)[kBvhWidth], int& ni,
                                             uint32_t& tris) {
  // every slot's entry distance uses the culling limit of the hit found
  // before the node, also after the first half's leaves have been tested
  const float limit = h.t * kCullSlack;
#pragma unroll
  for (int hf = 0; hf < kBvhHalves; ++hf) {
    float4 lx, hx, ly, hy, lz, hz;
    int4 ch;
    if (kScalar) {
      const UniformPtr np{S.nodes + node};
      lx = load4(np, 8 * hf + 0); hx = load4(np, 8 * hf + 1); ly = load4(np, 8 * hf + 2);
      hy = load4(np, 8 * hf + 3); lz = load4(np, 8 * hf + 4); hz = load4(np, 8 * hf + 5);
      ch = load_int4_uniform(np.p, 8 * hf + 6);
    } else {
      const float4* np = reinterpret_cast<const float4*>(S.nodes + node) + 8 * hf;
      lx = np[0]; hx = np[1]; ly = np[2]; hy = np[3]; lz = np[4]; hz = np[5];
      ch = reinterpret_cast<const int4*>(np)[6];
    }
    const float slo_x[kHalfWidth] = {lx.x, lx.y, lx.z, lx.w}, shi_x[kHalfWidth] = {hx.x, hx.y, hx.z, hx.w};
    const float slo_y[kHalfWidth] = {ly.x, ly.y, ly.z, ly.w}, shi_y[kHalfWidth] = {hy.x, hy.y, hy.z, hy.w};
    const float slo_z[kHalfWidth] = {lz.x, lz.y, lz.z, lz.w}, shi_z[kHalfWidth] = {hz.x, hz.y, hz.z, hz.w};
    const int32_t cc[kHalfWidth] = {ch.x, ch.y, ch.z, ch.w};
#pragma unroll
    for (int s = 0; s < kHalfWidth; ++s) {
      const int32_t c = cc[s];
      // the child word is wave-uniform: an empty slot is skipped on the
      // scalar unit before its slab test (the tree's empty slots are ~18 %)
      if (c == kChildEmpty) continue;
      const float te = slab_enter<kParallel>(rp, slo_x[s], shi_x[s], slo_y[s], shi_y[s], slo_z[s], shi_z[s], limit);
      const bool mine = active && te < INFINITY;
      if (__ballot(mine) == 0ull) continue;
      if (child_is_leaf(c)) {
        tris += leaf_count(c);
        // inactive lanes test a dummy ray whose result is discarded
        if (kScalar)
          test_leaf(UniformPtr{S.tris}, c, rp, h);
        else
          test_leaf_wave(S.tris, c, rp, h);
      } else {
        // order key: the entry distance of the first lane that enters the child
        const uint64_t m = __ballot(mine);
        const int lead = __builtin_amdgcn_readfirstlane(__ffsll((long long)m) - 1);
        const uint32_t key = (uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(te), lead);
        sorted_insert(ik, ic, ni, key, c);
        ++ni;
      }
    }
  }
}

// steps / tris count the packet's node visits and the triangles of the leaves
// it enters (wave-uniform: what the wave fetches from memory).
template <bool kParallel, bool kScalar>
__device__ __forceinline__ void closest_hit_packet_impl(const SceneView& S, const RayPre& rp, bool active, Hit& h,
                                                        uint32_t& steps, uint32_t& tris) {
  int sp = 0;
  int node = 0;
  int stk0 = 0, stk1 = 0;
  const int lane = (int)(threadIdx.x & 63);
  for (;;) {
    node = __builtin_amdgcn_readfirstlane(node);
    int ni = 0;
    int ic[kBvhWidth];
    uint32_t ik[kBvhWidth];
    ++steps;
    packet_visit<kParallel, kScalar>(S, rp, active, node, h, ic, ik, ni, tris);
    if (ni == 0) {
      if (sp == 0) break;
      --sp;
      node = stack_top(stk0, stk1, sp);
      continue;
    }
    // entered children are in near-to-far order: push all but the nearest,
    // farthest first, and visit the nearest
#pragma unroll
    for (int a = kBvhWidth - 1; a >= 1; --a) {
      if (a < ni) {
        stack_push(stk0, stk1, lane, sp, ic[a]);
        ++sp;
      }
    }
    node = ic[0];
  }
}

template <bool kScalar = false>
__device__ __forceinline__ void closest_hit_packet(const SceneView& S, const RayPre& rp, bool active, Hit& h,
                                                   uint32_t& steps, uint32_t& tris) {
  h.t = rp.tmax;
  h.b1 = h.b2 = 0.0f;
  h.k = 0xffffffffu;
  h.id = 0xffffffffu;
  if (S.node_count == 0) return;
  // The stack's entry k lives in lane k & 63, written by a lane select: a
  // lane outside EXEC would drop its entries.  Every caller enters with the
  // whole wave; a partial wave is flagged to the host, which discards the
  // launch's results, and returns misses instead of walking the tree.
  if (__ballot(1) != ~0ull) {
    if (S.fault) *S.fault = 1u;
    return;
  }
  if (__ballot(active && rp.par != 0) != 0ull)
    closest_hit_packet_impl<true, kScalar>(S, rp, active, h, steps, tris);
  else
    closest_hit_packet_impl<false, kScalar>(S, rp, active, h, steps, tris);
  finish_hit(h);
}

// The value of lane (lane ^ M), M a power of two below 64, without an
// address register: ds_swizzle in bit mode (and 0x1f, xor M) within each
// 32-lane half for M <= 16, and gfx950's v_permlane32_swap across the halves
// for M = 32.  __shfl_xor needs a (lane ^ M) * 4 address per mask, which the
// compiler computes once per kernel and keeps live (in the unit renderer, six
// of them spilled to scratch and reloaded in every round).
template <int M>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v) {
  static_assert(M >= 1 && M <= 32 && (M & (M - 1)) == 0, "power of two below 64");
  if constexpr (M == 32) {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (threadIdx.x & 32u) ? r[0] : r[1];
  } else {
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (M << 10) | 0x1f);
  }
}
template <int M>
__device__ __forceinline__ float lane_xor(float v) {
  return __uint_as_float(lane_xor<M>(__float_as_uint(v)));
}

__device__ __forceinline__ float wave_sum(float v) {
  v = v + lane_xor<32>(v);
  v = v + lane_xor<16>(v);
  v = v + lane_xor<8>(v);
  v = v + lane_xor<4>(v);
  v = v + lane_xor<2>(v);
  v = v + lane_xor<1>(v);
  return v;
}

__device__ __forceinline__ uint32_t unit_count(const UnitLaunch& U, uint64_t u, int i) {
  if (U.unit_counts) return U.unit_counts[u];
  if (U.omm_counts) return U.omm_counts[i];
  return (uint32_t)U.spp;
}

// Segmented form of wave_sum: lanes [a, a + seg) with a a multiple of seg
// (a power of two) sum their values; lane a ends with the total.  For a unit
// whose samples occupy lanes a .. a + n - 1 (n <= seg, other lanes 0) this is
// bit-identical to wave_sum over a wave holding the same samples in lanes
// 0 .. n - 1: the butterfly steps with offset >= seg only add zeros there.
template <int M>
__device__ __forceinline__ float segment_step(float v, int seg) {
  const float w = lane_xor<M>(v);
  return M < seg ? v + w : v;
}
__device__ __forceinline__ float segment_sum(float v, int seg) {
  v = segment_step<32>(v, seg);
  v = segment_step<16>(v, seg);
  v = segment_step<8>(v, seg);
  v = segment_step<4>(v, seg);
  v = segment_step<2>(v, seg);
  v = segment_step<1>(v, seg);
  return v;
}

// Pose, ommatidium and output slot of queue position u.
// The unit renderer's launches hold fewer than 2^32 units (launch_render_units
// checks), so its queue positions and output slots are 32-bit.
__device__ __forceinline__ void locate_unit32(const UnitLaunch& U, uint32_t u, int* p, int* i, uint32_t* out) {
  const uint32_t n = (uint32_t)U.omm_count;
  const uint32_t q = u / n;
  const uint32_t k = u - q * n;
  *p = (int)q;
  *i = U.order ? (int)U.order[k] : (int)k;
  *out = q * n + (uint32_t)*i;
}

__device__ __forceinline__ void locate_unit(const UnitLaunch& U, uint64_t u, int* p, int* i, uint64_t* out) {
  *p = (int)(u / (uint64_t)U.omm_count);
  const int k = (int)(u - (uint64_t)*p * (uint64_t)U.omm_count);
  *i = U.order ? (int)U.order[k] : k;
  *out = (uint64_t)*p * (uint64_t)U.omm_count + (uint64_t)*i;
}

// Bitonic sort of one value per lane, largest first, across the 64 lanes.
template <int K, int J>
__device__ __forceinline__ uint32_t sort_step(uint32_t v, int lane) {
  const uint32_t o = lane_xor<J>(v);
  const bool take_max = ((lane & J) == 0) == ((lane & K) == 0);
  return take_max ? (o > v ? o : v) : (o < v ? o : v);
}
template <int K, int J>
__device__ __forceinline__ uint32_t sort_merge(uint32_t v, int lane) {
  v = sort_step<K, J>(v, lane);
  if constexpr (J > 1) v = sort_merge<K, J / 2>(v, lane);
  return v;
}
__device__ __forceinline__ uint32_t wave_sort_desc(uint32_t v) {
  const int lane = (int)(threadIdx.x & 63);
  v = sort_merge<2, 1>(v, lane);
  v = sort_merge<4, 2>(v, lane);
  v = sort_merge<8, 4>(v, lane);
  v = sort_merge<16, 8>(v, lane);
  v = sort_merge<32, 16>(v, lane);
  v = sort_merge<64, 32>(v, lane);
  return v;
}

__device__ __forceinline__ int pow2_at_least(uint32_t n) {
  int s = 1;
  while ((uint32_t)s < n) s <<= 1;
  return s;
}

// Persistent unit renderer.  A wave takes units_per_grab consecutive units
// from the queue and renders them in rounds: a unit of more than 32 samples
// fills the wave alone (lane l takes samples l, l + 64, ...); smaller units are
// packed side by side into power-of-two lane segments, so the lanes of a
// wave stay busy when sample counts are small (adaptive pilots, uneven
// counts).  Each unit's colour is the same canonical sum either way.
//
// Packet variants: six waves per SIMD (at most 80 VGPRs).  Traversal is
// latency-bound, so resident waves set the throughput.  Left free, the
// compiler takes 120-138 VGPRs (three or four waves); under the bound it moves
// unit bookkeeping to scratch and reloads it outside the traversal loop.  The
// per-ray variant's LDS stack columns allow few resident blocks anyway; four
// waves (128 VGPRs) keep its traversal loop free of scratch.
// kOnePass: every unit has at most 64 samples (a uniform count), so a round's
// lanes trace once; the sums then need no registers across the traversal.
template <bool kStats, bool kSamples, int kTrav, bool kOnePass = false>
__global__ __launch_bounds__(kTraceThreads, kTrav == kUnitPerRay ? 4 : 6) void k_render_units(SceneView S,
                                                                                             UnitLaunch U) {
  // LDS (per-ray traversal only): one stack column per lane, laid out
  // [depth][blockDim.x]; packets keep their stack in VGPR lanes
  extern __shared__ uint32_t unit_lds[];
  const int lane = threadIdx.x & 63;
  uint32_t* lane_stack = unit_lds + threadIdx.x;
  const uint64_t units = (uint64_t)U.pose_count * (uint64_t)U.omm_count;
  // The queue is split into kXcds slices of consecutive units with a counter
  // each.  A wave takes units from its workgroup's XCD slice (neighbouring
  // directions of the same poses share that XCD's L2) and, once the slice is
  // empty, from the following slices; it exits when all are empty.
  const uint64_t slice = (units + kXcds - 1) / kXcds;
  uint32_t x = blockIdx.x % kXcds;
  uint32_t tried = 0;
  uint32_t w_steps = 0, w_tris = 0;   // packet node visits and leaf triangles of this wave
  for (;;) {
    uint64_t base = 0, s_end = 0;
    while (tried < kXcds) {
      const uint64_t s0 = (uint64_t)x * slice, s1 = min(units, s0 + slice);
      uint32_t off = 0xffffffffu;
      if (lane == 0 && s0 < s1) off = atomicAdd(U.queue + x, (uint32_t)U.units_per_grab);
      off = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(off, 0));
      if (s0 < s1 && s0 + off < s1) {
        base = s0 + off;
        s_end = s1;
        break;
      }
      x = (x + 1u) % kXcds;
      ++tried;
    }
    if (tried >= kXcds) {
      if (U.packet_counts && lane == 0 && w_steps != 0u) {
        atomicAdd(U.packet_counts, (unsigned long long)w_steps);
        atomicAdd(U.packet_counts + 1, (unsigned long long)w_tris);
      }
      break;
    }
    const int G = (int)min((uint64_t)U.units_per_grab, s_end - base);
    // lane g < G holds the sample count of unit base + g
    uint32_t my_n = 0;
    if (lane < G) {
      int p, i;
      uint32_t out;
      locate_unit32(U, (uint32_t)base + (uint32_t)lane, &p, &i, &out);
      my_n = unit_count(U, out, i);
    }
    // Rounds take the grab's units largest first: power-of-two segments in
    // decreasing size tile the 64 lanes without gaps.  Sorted position q
    // holds unit offset order[q] (units above 32 samples, which fill whole
    // rounds, compare equal).  Output positions stay per unit.
    const uint32_t key = wave_sort_desc(lane < G ? ((min(my_n, 64u) + 1u) << 6) | (uint32_t)lane : 0u);
    const int order = (int)(key & 63u);
    my_n = (uint32_t)__shfl((int)my_n, order);
    int g = 0;
    while (g < G) {
      // ---- build a round: which unit (slot) each lane serves ----
      int slot = -1, seg_off = 0, seg = 64;
      int g_end = g + 1;
      const uint32_t n_first = (uint32_t)__builtin_amdgcn_readlane((int)my_n, g);
      if (n_first > 32u) {
        slot = g;
      } else {
        int cursor = 0;
        g_end = g;
        while (g_end < G) {
          const uint32_t nj = (uint32_t)__builtin_amdgcn_readlane((int)my_n, g_end);
          if (nj > 32u) break;
          const int sj = pow2_at_least(nj);
          const int a = (cursor + sj - 1) & ~(sj - 1);
          if (a + sj > 64) break;
          if (lane >= a && lane < a + sj) {
            slot = g_end;
            seg_off = a;
            seg = sj;
          }
          cursor = a + sj;
          ++g_end;
        }
      }
      // ---- this lane's unit ----
      const int src = slot < 0 ? 0 : slot;
      const int unit_off = __shfl(order, src);
      int p, i;
      {
        uint32_t out_unused;
        locate_unit32(U, (uint32_t)base + (uint32_t)unit_off, &p, &i, &out_unused);
      }
      const uint32_t n_src = (uint32_t)__shfl((int)my_n, src);  // every lane takes part in the shuffle
      const uint32_t n = slot < 0 ? 0u : n_src;
      const float inv = 1.0f / (float)(n > 0u ? n : 1u);
      // The lane's place in the round, the one register of round bookkeeping
      // kept across the traversal (the rest is recomputed after it, which
      // keeps the six-wave kernel's unit state out of scratch): bit 16 set
      // when the lane serves a unit, bits 8-10 log2 of its segment, bits 0-6
      // the lane's first sample in its unit (lane - segment start).
      const uint32_t desc = (slot >= 0 ? 0x10000u : 0u) | ((uint32_t)__builtin_ctz((unsigned)seg) << 8) |
                            (uint32_t)(lane - seg_off);
      float r = 0.0f, gg = 0.0f, b = 0.0f;
      float sr = 0.0f, sg = 0.0f, sb = 0.0f, sd = 0.0f, sd2 = 0.0f, sh = 0.0f;
      for (uint32_t c0 = 0; __ballot(c0 + (desc & 127u) < n) != 0ull; c0 += 64u) {
        const uint32_t s = c0 + (desc & 127u);
        const bool active = s < n;
        Ray ray;
        if (active) {
          const U4 rnd = sample_randoms(U.seed, U.eye, (uint32_t)i, U.sample_base + s, U.frame,
                                        U.pose_base + (uint32_t)p);
          ray = generate_ray(U.omms[i], U.poses[p], rnd);
        } else {
          ray.ox = ray.oy = ray.oz = 0.0f;
          ray.dx = ray.dy = 0.0f;
          ray.dz = 1.0f;
          ray.tmin = 0.0f;
          ray.tmax = 1.0f;
        }
        const RayPre rp = precompute_ray(ray);
        Hit h;
        if (kTrav == kUnitPerRay) {
          h.t = rp.tmax;
          h.k = h.id = 0xffffffffu;
          h.b1 = h.b2 = 0.0f;
          if (active) closest_hit(S, rp, lane_stack, (int)blockDim.x, h);
        } else {
          closest_hit_packet<kTrav == kUnitPacketScalar>(S, rp, active, h, w_steps, w_tris);
        }
        if (active) {
          const V3 c = shade(S, h, ray.dx, ray.dy, ray.dz);
          r = r + c.x * inv;
          gg = gg + c.y * inv;
          b = b + c.z * inv;
          const bool hit = h.id != 0xffffffffu;
          if (kSamples)
            U.samples[U.sample_offsets[(uint32_t)p * (uint32_t)U.omm_count + (uint32_t)i] + s] =
                make_float4(c.x, c.y, c.z, hit ? h.t : -1.0f);
          if (kStats) {
            sr += c.x * c.x;
            sg += c.y * c.y;
            sb += c.z * c.z;
            const float idp = hit ? 1.0f / (1.0f + h.t) : 0.0f;
            sd += idp;
            sd2 += idp * idp;
            sh += hit ? 1.0f : 0.0f;
          }
        }
        if (kOnePass) break;
      }
      const int seg_len = 1 << ((desc >> 8) & 7u);
      r = segment_sum(r, seg_len);
      gg = segment_sum(gg, seg_len);
      b = segment_sum(b, seg_len);
      if (kStats) {
        sr = segment_sum(sr, seg_len);
        sg = segment_sum(sg, seg_len);
        sb = segment_sum(sb, seg_len);
        sd = segment_sum(sd, seg_len);
        sd2 = segment_sum(sd2, seg_len);
        sh = segment_sum(sh, seg_len);
      }
      if ((desc & 0x1007fu) == 0x10000u) {   // the first lane of a unit's segment
        const uint32_t out = (uint32_t)p * (uint32_t)U.omm_count + (uint32_t)i;
        U.colours[out] = make_float4(r, gg, b, (float)n);
        if (kStats) {
          // E[c^2] - E[c]^2 with E[c] the unit's mean colour
          const float vr = fmaxf(sr * inv - r * r, 0.0f);
          const float vg = fmaxf(sg * inv - gg * gg, 0.0f);
          const float vb = fmaxf(sb * inv - b * b, 0.0f);
          const float md = sd * inv;
          U.stats[2 * (size_t)out] = make_float4(vr, vg, vb, sh * inv);
          U.stats[2 * (size_t)out + 1] = make_float4(md, fmaxf(sd2 * inv - md * md, 0.0f), 0.0f, 0.0f);
        }
      }
      g = g_end;
    }
  }
}

// Workgroups are handed to the 8 XCDs round-robin (block b to XCD b % 8),
// and each XCD has its own L2.  Logical blocks are therefore dealt to XCDs in
// runs of kXcdRun consecutive blocks: the queue is in Morton order of the
// ommatidial axes, so an XCD's resident waves look in neighbouring directions
// and share BVH nodes in its L2, while runs still alternate between XCDs so
// that each gets its share of costly directions.  The mapping is a
// bijection, so results do not depend on it.
constexpr uint32_t kXcdRun = 16;

__device__ __forceinline__ uint32_t xcd_logical_block(uint32_t b, uint32_t blocks) {
  const uint32_t span = kXcds * kXcdRun;
  if (b >= (blocks / span) * span) return b;
  const uint32_t x = b % kXcds, j = b / kXcds;
  return ((j / kXcdRun) * kXcds + x) * kXcdRun + j % kXcdRun;
}

// Split launch of units with a uniform sample count n, per-ray traversal:
// task t is chunk t % C of queue position t / C (C = ceil(n / 64)), so that a
// single pose's few units still fill the GPU.  Wave w of logical block L is
// task 4L + w; there is no task loop.  Lane l of a task keeps the terms of
// its one sample (scaled colour, and the statistics terms) in `partial`; the
// wave that completes a unit's last chunk adds the chunks lane by lane in
// chunk order and then across the wave, which gives the same sums as
// k_render_units (there a lane adds samples l, l + 64, ... in this order).
//
// The rays of a task are traversed one per lane (an LDS stack column per
// lane), so lanes past n simply skip the traversal.
//
// Why no persistent task loop: a loop whose body takes its task from a lane-0
// atomic broadcast with __shfl and ends in lane-0-only stores is compiled
// with the lane-0 branch leaving the loop on its own; the other 63 lanes
// then go round without lane 0, read 0 from the broadcast (a permute from a
// disabled lane) and, for C = 1, never leave.  Every wave here runs
// straight through, and its lane-dependent branches rejoin before the next
// cross-lane operation.
template <bool kStats, bool kSamples>
__global__ __launch_bounds__(kTraceThreads, 5) void k_render_chunks(SceneView S, UnitLaunch U, int chunks) {
  extern __shared__ uint32_t lane_stacks[];
  const int lane = threadIdx.x & 63;
  uint32_t* lane_stack = lane_stacks + threadIdx.x;
  constexpr int kPer = kStats ? 3 : 1;
  const uint64_t units = (uint64_t)U.pose_count * (uint64_t)U.omm_count;
  const uint64_t tasks = units * (uint64_t)chunks;
  const uint64_t t = (uint64_t)xcd_logical_block(blockIdx.x, gridDim.x) * (kTraceThreads / 64) +
                     (uint64_t)(threadIdx.x >> 6);
  if (t >= tasks) return;  // whole wave
  const uint32_t n = (uint32_t)U.spp;
  const float inv = 1.0f / (float)n;
  const uint64_t uq = t / (uint64_t)chunks;
  const uint32_t j = (uint32_t)(t - uq * (uint64_t)chunks);
  int p, i;
  uint64_t out;
  locate_unit(U, uq, &p, &i, &out);
  const uint32_t s = j * 64u + (uint32_t)lane;
  const bool active = s < n;
  float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b = a, c2 = a;
  if (active) {
    const U4 rnd = sample_randoms(U.seed, U.eye, (uint32_t)i, U.sample_base + s, U.frame,
                                  U.pose_base + (uint32_t)p);
    const Ray ray = generate_ray(U.omms[i], U.poses[p], rnd);
    const RayPre rp = precompute_ray(ray);
    Hit h;
    closest_hit(S, rp, lane_stack, (int)blockDim.x, h);
    const V3 c = shade(S, h, ray.dx, ray.dy, ray.dz);
    const bool hit = h.id != 0xffffffffu;
    a = make_float4(c.x * inv, c.y * inv, c.z * inv, 0.0f);
    if (kSamples) U.samples[U.sample_offsets[out] + s] = make_float4(c.x, c.y, c.z, hit ? h.t : -1.0f);
    if (kStats) {
      const float idp = hit ? 1.0f / (1.0f + h.t) : 0.0f;
      b = make_float4(c.x * c.x, c.y * c.y, c.z * c.z, idp);
      c2 = make_float4(idp * idp, hit ? 1.0f : 0.0f, 0.0f, 0.0f);
    }
  }
  float4* slot = U.chunk_partial + ((uq * (uint64_t)chunks + j) * 64u + (uint64_t)lane) * kPer;
  slot[0] = a;
  if (kStats) {
    slot[1] = b;
    slot[2] = c2;
  }
  __threadfence();
  uint32_t prev = 0;
  if (lane == 0) prev = atomicAdd(U.chunk_done + uq, 1u);
  prev = __builtin_amdgcn_readfirstlane(__shfl(prev, 0));
  if (prev != (uint32_t)chunks - 1u) return;  // whole wave
  __threadfence();
  float r = 0.0f, gg = 0.0f, bb = 0.0f;
  float sr = 0.0f, sg = 0.0f, sb = 0.0f, sd = 0.0f, sd2 = 0.0f, sh = 0.0f;
  for (int q = 0; q < chunks; ++q) {
    if (q * 64 + lane < (int)n) {
      const float4* src = U.chunk_partial + ((uq * (uint64_t)chunks + (uint64_t)q) * 64u + (uint64_t)lane) * kPer;
      const float4 v = src[0];
      r = r + v.x;
      gg = gg + v.y;
      bb = bb + v.z;
      if (kStats) {
        const float4 w = src[1], x = src[2];
        sr += w.x;
        sg += w.y;
        sb += w.z;
        sd += w.w;
        sd2 += x.x;
        sh += x.y;
      }
    }
  }
  r = segment_sum(r, 64);
  gg = segment_sum(gg, 64);
  bb = segment_sum(bb, 64);
  if (kStats) {
    sr = segment_sum(sr, 64);
    sg = segment_sum(sg, 64);
    sb = segment_sum(sb, 64);
    sd = segment_sum(sd, 64);
    sd2 = segment_sum(sd2, 64);
    sh = segment_sum(sh, 64);
  }
  if (lane == 0) {
    U.colours[out] = make_float4(r, gg, bb, (float)n);
    if (kStats) {
      const float vr = fmaxf(sr * inv - r * r, 0.0f);
      const float vg = fmaxf(sg * inv - gg * gg, 0.0f);
      const float vb = fmaxf(sb * inv - bb * bb, 0.0f);
      const float md = sd * inv;
      U.stats[2 * out] = make_float4(vr, vg, vb, sh * inv);
      U.stats[2 * out + 1] = make_float4(md, fmaxf(sd2 * inv - md * md, 0.0f), 0.0f, 0.0f);
    }
  }
}

// The same split launch for wave-packet traversal: the whole wave returns
// together when past the last task, and nothing but the traversal sits
// between the ray generation and the chunk store, which keeps register use
// low (more resident waves).  Same per-lane chunk values, completion
// counters and last-wave reduction as k_render_chunks (no statistics).
template <bool kSamples, bool kScalar, bool kSplit>
__global__ __launch_bounds__(kTraceThreads, 6) void k_render_tasks(SceneView S, UnitLaunch U, int chunks) {
  const int lane = threadIdx.x & 63;
  const uint64_t units = (uint64_t)U.pose_count * (uint64_t)U.omm_count;
  const uint64_t tasks = units * (uint64_t)chunks;
  // With vector node loads the task, unit and output indices go to scalar
  // registers (the wave index read once as uniform), which keeps the split
  // variant from spilling; the scalar-load traversal needs those registers
  // for node words and keeps the indices in vector registers.
  const uint32_t wave = kScalar ? (threadIdx.x >> 6) : (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint64_t wv = (uint64_t)xcd_logical_block(blockIdx.x, gridDim.x) * (kTraceThreads / 64) + (uint64_t)wave;
  // With a long-task split the first kSplitSubWaves waves trace the listed
  // tasks: 4 waves per tier-0 task (wave w: part w % 4 of tier-0 task w / 4),
  // then 2 per tier-1 task, then one per tier-2 task (whole); the rest every
  // other task in order.
  const bool split = kSplit && U.split_flag != nullptr;
  const uint64_t sub_waves = split ? (uint64_t)kSplitSubWaves : 0u;
  uint64_t t;
  int part = -1, ways = 1;
  uint32_t li = 0;   // completion counter of a split task
  if (wv < sub_waves) {
    constexpr uint32_t r0 = 4u * kSplit4Max, r1 = r0 + 2u * kSplit2Max;
    const uint32_t w = (uint32_t)wv;
    uint32_t tier, k, first;
    if (w < r0) {
      tier = 0u;
      k = w >> 2;
      part = (int)(w & 3u);
      ways = 4;
      first = 0u;
    } else if (w < r1) {
      tier = 1u;
      k = (w - r0) >> 1;
      part = (int)((w - r0) & 1u);
      ways = 2;
      first = kSplit4Max;
    } else {
      tier = 2u;
      k = w - r1;
      first = kSplit4Max + kSplit2Max;
    }
    const uint32_t listed = (uint32_t)__builtin_amdgcn_readfirstlane((int)U.split_count[tier]);
    if (k >= listed) return;  // whole wave
    t = (uint32_t)__builtin_amdgcn_readfirstlane((int)U.split_list[first + k]);
    li = first + k;
  } else {
    t = wv - sub_waves;
    if (t >= tasks) return;  // whole wave
    if (split && __builtin_amdgcn_readfirstlane((int)U.split_flag[t]) != 0) return;  // traced listed
  }
  const uint32_t n = (uint32_t)U.spp;
  const float inv = 1.0f / (float)n;
  const uint64_t uq = t / (uint64_t)chunks;
  const uint32_t j = (uint32_t)(t - uq * (uint64_t)chunks);
  int p, i;
  uint64_t out;
  locate_unit(U, uq, &p, &i, &out);
  const uint32_t s = j * 64u + (uint32_t)lane;
  const bool active = s < n;
  Ray ray;
  if (active) {
    const U4 rnd = sample_randoms(U.seed, U.eye, (uint32_t)i, U.sample_base + s, U.frame,
                                  U.pose_base + (uint32_t)p);
    ray = generate_ray(U.omms[i], U.poses[p], rnd);
  } else {
    ray.ox = ray.oy = ray.oz = 0.0f;
    ray.dx = ray.dy = 0.0f;
    ray.dz = 1.0f;
    ray.tmin = 0.0f;
    ray.tmax = 1.0f;
  }
  const RayPre rp = precompute_ray(ray);
  // A split packet traces the lanes whose rays rank part * 64 / ways ..
  // (part + 1) * 64 / ways - 1 by elevation (the direction's y, ties by
  // lane): rays of neighbouring elevation fan out less than the whole chunk.
  bool mine = active;
  if (part >= 0) {
    const uint32_t bits = __float_as_uint(ray.dy);
    const uint32_t key = ((bits & 0x80000000u) ? ~bits : (bits | 0x80000000u)) & ~63u;
    const uint32_t sorted = wave_sort_desc(key | (uint32_t)lane);
    // sorted position q holds lane (sorted & 63): send q back to that lane
    const int rank = __builtin_amdgcn_ds_permute((int)((sorted & 63u) * 4u), lane);
    mine = active && (rank >> (ways == 4 ? 4 : 5)) == part;
  }
  Hit h;
  uint32_t w_steps = 0, w_tris = 0;
  closest_hit_packet<kScalar>(S, rp, mine, h, w_steps, w_tris);
  if (U.packet_counts && lane == 0) {
    atomicAdd(U.packet_counts, (unsigned long long)w_steps);
    atomicAdd(U.packet_counts + 1, (unsigned long long)w_tris);
  }
  if (kSplit && U.split_cost && lane == 0) atomicAdd(U.split_cost + t, w_steps);
  float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (mine) {
    const V3 c = shade(S, h, ray.dx, ray.dy, ray.dz);
    a = make_float4(c.x * inv, c.y * inv, c.z * inv, 0.0f);
    if (kSamples)
      U.samples[U.sample_offsets[out] + s] = make_float4(c.x, c.y, c.z, h.id != 0xffffffffu ? h.t : -1.0f);
  }
  // a split packet writes its own lanes' terms; the other packets of its
  // task write the rest
  if (part < 0 || mine) U.chunk_partial[(uq * (uint64_t)chunks + j) * 64u + (uint64_t)lane] = a;
  __threadfence();
  if (part >= 0) {
    // the last of the task's split packets to finish completes the task
    uint32_t done = 0;
    if (lane == 0) done = atomicAdd(U.split_sub_done + li, 1u);
    done = __builtin_amdgcn_readfirstlane(__shfl(done, 0));
    if (done != (uint32_t)ways - 1u) return;  // whole wave
    __threadfence();
  }
  uint32_t prev = 0;
  if (lane == 0) prev = atomicAdd(U.chunk_done + uq, 1u);
  prev = __builtin_amdgcn_readfirstlane(__shfl(prev, 0));
  if (prev != (uint32_t)chunks - 1u) return;  // whole wave
  __threadfence();
  float r = 0.0f, gg = 0.0f, bb = 0.0f;
  for (int q = 0; q < chunks; ++q) {
    if (q * 64 + lane < (int)n) {
      const float4 v = U.chunk_partial[(uq * (uint64_t)chunks + (uint64_t)q) * 64u + (uint64_t)lane];
      r = r + v.x;
      gg = gg + v.y;
      bb = bb + v.z;
    }
  }
  r = segment_sum(r, 64);
  gg = segment_sum(gg, 64);
  bb = segment_sum(bb, 64);
  if (lane == 0) U.colours[out] = make_float4(r, gg, bb, (float)n);
}

// Failure-handling probe of the packet traversal's partial-wave guard: one
// wave traces 64 rays, lanes at or above active_lanes skip the call (their
// ids stay 0xfffffffe).
__global__ __launch_bounds__(64) void k_packet_guard_probe(SceneView S, const Ray* __restrict__ rays,
                                                           int active_lanes, int32_t* __restrict__ ids) {
  const int lane = (int)(threadIdx.x & 63);
  Hit h;
  h.id = 0xfffffffeu;
  if (lane < active_lanes) {
    const RayPre rp = precompute_ray(rays[lane]);
    uint32_t steps = 0, tris = 0;
    closest_hit_packet<false>(S, rp, true, h, steps, tris);
  }
  ids[lane] = (int32_t)h.id;
}

// Long-task split plan, one workgroup launched before the task kernel: it
// clears the units' completion counters, takes a histogram of the last
// launch's per-task packet steps (bins of 8 steps, the last one open), and
// picks each tier's threshold: the smallest multiple of 8 steps at or above
// the tier's minimum with at most the tiers' combined room (tier 0: 128
// tasks, tiers 0-1: 1,024, tiers 0-2: 3,072) at or above it (none when even
// the open top bin holds more).  It lists the tasks by tier and flags them,
// and clears the split packets' completion counters and every task's step
// count for this launch to fill in.
constexpr int kSplitPlanThreads = 1024;
constexpr int kSplitPlanPer = 16;   // step counts a thread holds at once
static_assert(kSplitBins <= kSplitPlanThreads, "one thread per histogram bin");

__device__ __forceinline__ uint32_t split_threshold(uint32_t over, uint32_t min_steps) {
  const uint32_t b = max(over, min_steps / 8u);
  return b >= (uint32_t)kSplitBins ? 0xffffffffu : b * 8u;
}

__global__ __launch_bounds__(kSplitPlanThreads) void k_split_plan(uint32_t* __restrict__ cost, uint64_t tasks,
                                                                   uint32_t* __restrict__ flag,
                                                                   uint32_t* __restrict__ list,
                                                                   uint32_t* __restrict__ count,
                                                                   uint32_t* __restrict__ sub_done,
                                                                   uint32_t* __restrict__ unit_done, uint64_t units) {
  __shared__ uint32_t hist[kSplitBins];
  __shared__ uint32_t listed[3];
  const uint32_t tid = threadIdx.x;
  for (uint32_t b = tid; b < (uint32_t)kSplitBins; b += kSplitPlanThreads) hist[b] = 0u;
  if (tid < 3u) listed[tid] = 0u;
  for (uint32_t i = tid; i < (uint32_t)kSplitSubMax; i += kSplitPlanThreads) sub_done[i] = 0u;
  for (uint64_t u = tid; u < units; u += kSplitPlanThreads) unit_done[u] = 0u;
  __syncthreads();
  // a round's loads are all issued before its histogram updates; only tasks
  // of at least kSplitEarlyMinSteps enter the histogram (shorter ones are
  // never listed, and the threshold searches start there or above)
  const uint64_t per_round = (uint64_t)kSplitPlanThreads * kSplitPlanPer;
  uint32_t c[kSplitPlanPer];
  for (uint64_t base = 0; base < tasks; base += per_round) {
#pragma unroll
    for (int k = 0; k < kSplitPlanPer; ++k) {
      const uint64_t t = base + (uint64_t)k * kSplitPlanThreads + tid;
      c[k] = t < tasks ? cost[t] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kSplitPlanPer; ++k)
      if (c[k] >= kSplitEarlyMinSteps) atomicAdd(&hist[min(c[k] / 8u, (uint32_t)kSplitBins - 1u)], 1u);
  }
  __syncthreads();
  // suffix counts S[b] = tasks in bins b .. kSplitBins - 1 fall as b grows,
  // so the bins with S[b] above a room are a prefix of the bins
  for (uint32_t off = 1; off < (uint32_t)kSplitBins; off <<= 1) {
    const uint32_t v = tid + off < (uint32_t)kSplitBins ? hist[tid + off] : 0u;
    __syncthreads();
    if (tid < (uint32_t)kSplitBins) hist[tid] += v;
    __syncthreads();
  }
  const bool bin = tid < (uint32_t)kSplitBins;
  const uint32_t over0 = (uint32_t)__syncthreads_count(bin && hist[tid] > (uint32_t)kSplit4Max);
  const uint32_t over1 = (uint32_t)__syncthreads_count(bin && hist[tid] > (uint32_t)(kSplit4Max + kSplit2Max));
  const uint32_t over2 = (uint32_t)__syncthreads_count(bin && hist[tid] > (uint32_t)kSplitListMax);
  const uint32_t th0 = split_threshold(over0, kSplitMinSteps);
  const uint32_t th1 = split_threshold(over1, kSplitMinSteps);
  const uint32_t th2 = split_threshold(over2, kSplitEarlyMinSteps);
  for (uint64_t base = 0; base < tasks; base += per_round) {
    if (tasks > per_round) {   // one round: the step counts are still held
#pragma unroll
      for (int k = 0; k < kSplitPlanPer; ++k) {
        const uint64_t t = base + (uint64_t)k * kSplitPlanThreads + tid;
        c[k] = t < tasks ? cost[t] : 0u;
      }
    }
#pragma unroll
    for (int k = 0; k < kSplitPlanPer; ++k) {
      const uint64_t t = base + (uint64_t)k * kSplitPlanThreads + tid;
      if (t < tasks) {
        uint32_t f = 0u;
        const uint32_t tier = c[k] >= th0 ? 0u : c[k] >= th1 ? 1u : c[k] >= th2 ? 2u : 3u;
        if (tier < 3u) {
          const uint32_t room = tier == 0u ? kSplit4Max : tier == 1u ? kSplit2Max : kSplit1Max;
          const uint32_t first = tier == 0u ? 0u : tier == 1u ? kSplit4Max : kSplit4Max + kSplit2Max;
          const uint32_t slot = atomicAdd(&listed[tier], 1u);
          if (slot < room) {
            list[first + slot] = (uint32_t)t;
            f = tier + 1u;
          }
        }
        flag[t] = f;
        cost[t] = 0u;
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    count[0] = min(listed[0], (uint32_t)kSplit4Max);
    count[1] = min(listed[1], (uint32_t)kSplit2Max);
    count[2] = min(listed[2], (uint32_t)kSplit1Max);
    count[3] = th0;
    count[4] = th1;
    count[5] = th2;
  }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Traversal work counters of the same rays (nodes visited, triangles tested),
// summed over the launch: the algorithmic memory traffic of the trace kernel.
__global__ __launch_bounds__(kTraceThreads) void k_trace_stats(SceneView S, EyeLaunch L, uint64_t total,
                                                               unsigned long long* __restrict__ totals) {
  extern __shared__ uint32_t lds_stack[];
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t nv = 0, nt = 0, hit = 0;
  if (r < total) {
    int p, i, s, n;
    locate_ray(L, r, &p, &i, &s, &n);
    const U4 rnd = sample_randoms(L.seed, L.eye, (uint32_t)i, (uint32_t)s, L.frame, L.pose_base + (uint32_t)p);
    const Ray ray = generate_ray(L.omms[i], L.poses[p], rnd);
    Hit h;
    closest_hit<true>(S, precompute_ray(ray), lds_stack + threadIdx.x, blockDim.x, h, &nv, &nt);
    hit = h.id != 0xffffffffu ? 1u : 0u;
  }
  unsigned long long a = nv, b = nt, c = hit;
  for (int m = 32; m >= 1; m >>= 1) {
    a += __shfl_xor(a, m);
    b += __shfl_xor(b, m);
    c += __shfl_xor(c, m);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&totals[0], a);
    atomicAdd(&totals[1], b);
    atomicAdd(&totals[2], c);
  }
}

__global__ void k_generate_rays(EyeLaunch L, uint64_t total, Ray* __restrict__ rays) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= total) return;
  int p, i, s, n;
  locate_ray(L, r, &p, &i, &s, &n);
  const U4 rnd = sample_randoms(L.seed, L.eye, (uint32_t)i, (uint32_t)s, L.frame, L.pose_base + (uint32_t)p);
  rays[r] = generate_ray(L.omms[i], L.poses[p], rnd);
}

__global__ __launch_bounds__(kTraceThreads) void k_trace_rays(SceneView S, const Ray* __restrict__ rays,
                                                              uint64_t n, int32_t* __restrict__ hit,
                                                              float4* __restrict__ tbb) {
  extern __shared__ uint32_t lds_stack[];
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const RayPre rp = precompute_ray(rays[r]);
  Hit h;
  closest_hit(S, rp, lds_stack + threadIdx.x, blockDim.x, h);
  hit[r] = h.id == 0xffffffffu ? -1 : (int32_t)h.id;
  tbb[r] = make_float4(h.t, h.b1, h.b2, 0.0f);
}

__device__ __forceinline__ uchar4 make_color(float4 c) {
  return make_uchar4(encode_channel(c.x), encode_channel(c.y), encode_channel(c.z), 255);
}

// One pixel of a regular camera per thread: primary ray, closest hit
// (per-ray LDS stack), shading or sky, make_color.  rays (nullable) receives
// the generated rays.
__global__ __launch_bounds__(kTraceThreads) void k_camera_frame(SceneView S, CameraLaunch C,
                                                                uchar4* __restrict__ frame, Ray* __restrict__ rays) {
  extern __shared__ uint32_t lds_stack[];
  const int pixels = C.width * C.height;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pixels) return;
  const Ray r = camera_ray(C, p % C.width, p / C.width);
  if (rays) rays[p] = r;
  if (!frame) return;
  Hit h;
  closest_hit(S, precompute_ray(r), lds_stack + threadIdx.x, blockDim.x, h);
  const V3 c = shade(S, h, r.dx, r.dy, r.dz);
  frame[p] = make_color(make_float4(c.x, c.y, c.z, 0.0f));
}

__global__ void k_frame_sdf(const float4* __restrict__ colours, int omm_count, uchar4* __restrict__ frame,
                            int width) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= width || x >= omm_count) return;
  frame[x] = make_color(colours[x]);
}

__global__ void k_frame_sd(const float4* __restrict__ colours, int omm_count, uchar4* __restrict__ frame,
                           int width, int height) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= width || y >= height) return;
  const int i = (int)(((long long)x * omm_count) / width);
  frame[(size_t)y * width + x] = make_color(colours[i]);
}

// Column x shows ommatidium x's samples, row y its sample y.  offsets
// (nullable: spp samples each) is the per-ommatidium sample counts' prefix
// sum; rows at or beyond an ommatidium's count keep the frame's contents.
__global__ void k_frame_raw(const float4* __restrict__ samples, const uint32_t* __restrict__ offsets,
                            int omm_count, int spp, uchar4* __restrict__ frame, int width, int height) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= width || y >= height || x >= omm_count) return;
  const size_t first = offsets ? (size_t)offsets[x] : (size_t)x * (size_t)spp;
  const uint32_t n = offsets ? offsets[x + 1] - offsets[x] : (uint32_t)spp;
  if ((uint32_t)y >= n) return;
  // the reference stores c / spp per sample and displays it unsummed
  const float4 v = samples[first + (size_t)y];
  const float inv = 1.0f / (float)n;
  frame[(size_t)y * width + x] = make_color(make_float4(v.x * inv, v.y * inv, v.z * inv, 0.0f));
}

__global__ void k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

bool check_launch(const char* what, std::string* err) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    if (err) *err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
  }
  return true;
}

size_t stack_bytes(const DeviceScene& s) {
  const int depth = s.stack_bound > 0 ? s.stack_bound : 1;
  return sizeof(uint32_t) * (size_t)depth * kTraceThreads;
}

}  // namespace

static int g_trace_mode = kTracePacketScalar;
static const char* g_last_kernel = "";
static bool g_split = true;

template <int kTrav>
static void launch_units_variant(const UnitLaunch& U, const SceneView& v, int blocks, size_t lds,
                                 hipStream_t stream) {
  const dim3 grid(blocks), block(kTraceThreads);
  // a uniform count of at most 64 samples per unit (the adaptive pilot, uniform
  // batches of few samples): every round traces once
  const bool one = !U.omm_counts && !U.unit_counts && U.spp <= 64;
  if (U.stats && U.samples && one)
    hipLaunchKernelGGL((k_render_units<true, true, kTrav, true>), grid, block, lds, stream, v, U);
  else if (U.stats && one)
    hipLaunchKernelGGL((k_render_units<true, false, kTrav, true>), grid, block, lds, stream, v, U);
  else if (U.samples && one)
    hipLaunchKernelGGL((k_render_units<false, true, kTrav, true>), grid, block, lds, stream, v, U);
  else if (one)
    hipLaunchKernelGGL((k_render_units<false, false, kTrav, true>), grid, block, lds, stream, v, U);
  else if (U.stats && U.samples)
    hipLaunchKernelGGL((k_render_units<true, true, kTrav>), grid, block, lds, stream, v, U);
  else if (U.stats)
    hipLaunchKernelGGL((k_render_units<true, false, kTrav>), grid, block, lds, stream, v, U);
  else if (U.samples)
    hipLaunchKernelGGL((k_render_units<false, true, kTrav>), grid, block, lds, stream, v, U);
  else
    hipLaunchKernelGGL((k_render_units<false, false, kTrav>), grid, block, lds, stream, v, U);
}

static void launch_chunks(const UnitLaunch& U, const SceneView& v, int blocks, size_t lds, int chunks,
                          hipStream_t stream) {
  const dim3 grid(blocks), block(kTraceThreads);
  if (U.stats && U.samples)
    hipLaunchKernelGGL((k_render_chunks<true, true>), grid, block, lds, stream, v, U, chunks);
  else if (U.stats)
    hipLaunchKernelGGL((k_render_chunks<true, false>), grid, block, lds, stream, v, U, chunks);
  else if (U.samples)
    hipLaunchKernelGGL((k_render_chunks<false, true>), grid, block, lds, stream, v, U, chunks);
  else
    hipLaunchKernelGGL((k_render_chunks<false, false>), grid, block, lds, stream, v, U, chunks);
}

template <int kTrav>
static const void* units_kernel_variant(const UnitLaunch& U) {
  const bool one = !U.omm_counts && !U.unit_counts && U.spp <= 64;
  if (U.stats && U.samples && one) return reinterpret_cast<const void*>(&k_render_units<true, true, kTrav, true>);
  if (U.stats && one) return reinterpret_cast<const void*>(&k_render_units<true, false, kTrav, true>);
  if (U.samples && one) return reinterpret_cast<const void*>(&k_render_units<false, true, kTrav, true>);
  if (one) return reinterpret_cast<const void*>(&k_render_units<false, false, kTrav, true>);
  if (U.stats && U.samples) return reinterpret_cast<const void*>(&k_render_units<true, true, kTrav>);
  if (U.stats) return reinterpret_cast<const void*>(&k_render_units<true, false, kTrav>);
  if (U.samples) return reinterpret_cast<const void*>(&k_render_units<false, true, kTrav>);
  return reinterpret_cast<const void*>(&k_render_units<false, false, kTrav>);
}

static int device_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  return cus;
}

static size_t units_lds_bytes(const DeviceScene& scene, int traversal) {
  return traversal == kUnitPerRay ? stack_bytes(scene) : 0;
}

// Packets keep their traversal stack in VGPR lanes (kPacketStackMax entries);
// a deeper BVH takes the per-ray traversal, whose stacks live in LDS.
static int effective_traversal(const DeviceScene& scene, int traversal) {
  return traversal != kUnitPerRay && scene.stack_bound > kPacketStackMax ? (int)kUnitPerRay : traversal;
}

static const void* units_kernel_of(const UnitLaunch& U) {
  if (U.traversal == kUnitPerRay) return units_kernel_variant<kUnitPerRay>(U);
  if (U.traversal == kUnitPacketScalar) return units_kernel_variant<kUnitPacketScalar>(U);
  return units_kernel_variant<kUnitPacket>(U);
}

int render_units_capacity(const DeviceScene& scene, const UnitLaunch& U_in) {
  UnitLaunch U = U_in;
  U.traversal = effective_traversal(scene, U.traversal);
  const void* fn = units_kernel_of(U);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kTraceThreads, units_lds_bytes(scene, U.traversal)) !=
          hipSuccess ||
      per_cu <= 0)
    per_cu = 4;
  return device_cus() * per_cu;
}

bool launch_render_units(const DeviceScene& scene, const UnitLaunch& U_in, hipStream_t stream, std::string* err) {
  UnitLaunch U = U_in;
  const uint64_t units = (uint64_t)U.pose_count * (uint64_t)U.omm_count;
  if (units == 0) return true;
  if (units > 0xffffffffull - 65536ull) {
    if (err) *err = "render_units: too many units for one launch";
    return false;
  }
  if (!U.queue || !U.colours || (U.samples && !U.sample_offsets)) {
    if (err) *err = "render_units: incomplete launch description";
    return false;
  }
  if (U.traversal < kUnitPacket || U.traversal > kUnitPerRay) {
    if (err) *err = "render_units: unknown traversal";
    return false;
  }
  U.traversal = effective_traversal(scene, U.traversal);
  const size_t lds = units_lds_bytes(scene, U.traversal);
  if (lds > 160 * 1024) {
    if (err) *err = "render_units: BVH too deep for the LDS traversal stack";
    return false;
  }
  const int capacity = render_units_capacity(scene, U);
  int blocks = U.blocks > 0 ? U.blocks : capacity;
  const SceneView v = view_of(scene);
  // Split launch: uniform counts and fewer units than two per wave slot of
  // the device (a single pose, typically), when the caller gave scratch: each
  // unit's 64-sample chunks are tasks of their own, one wave each.  With
  // per-ray traversal or statistics, units of 32 samples or fewer stay in the
  // unit kernel, which packs two or more of them into a wave's lanes (a
  // per-ray task would leave the other lanes idle); the packet task kernel
  // takes any count (a single pose still renders sooner that way).  The slots
  // are counted at the kernels' launch bound (5 blocks of 4 waves per CU) or
  // the unit kernel's occupancy if higher, not lowered by the LDS the per-ray
  // stack takes: a deep tree's per-ray stack cuts the persistent unit kernel
  // to two blocks per CU, which leaves a single pose badly balanced.
  const uint64_t chunks = U.spp > 64 ? ((uint64_t)U.spp + 63u) / 64u : 1u;
  const uint64_t resident_waves = (uint64_t)std::max(capacity, device_cus() * 5) * (kTraceThreads / 64);
  const uint64_t per_chunk = U.stats ? 3u : 1u;
  const bool packet_tasks = U.traversal != kUnitPerRay && !U.stats;
  if (g_split && !U.omm_counts && !U.unit_counts && (U.spp > 32 || packet_tasks) && units < 2 * resident_waves &&
      U.chunk_partial && U.chunk_done && units * chunks <= 0xffffffffull - 65536ull &&
      units * chunks * 64u * per_chunk <= U.chunk_partial_cap && units <= U.chunk_done_cap) {
    const uint64_t tasks = units * chunks;
    if (packet_tasks && tasks / (kTraceThreads / 64) < 0x7fffffffull - (uint64_t)kSplitSubWaves) {
      // wave packets: the non-persistent task kernel, one wave per chunk
      UnitLaunch T = U;
      const bool have_split = U.split_cost && U.split_flag && U.split_list && U.split_count && U.split_sub_done &&
                              U.split_tasks == tasks;
      if (have_split && U.split_plan) {
        // the plan kernel also clears the units' completion counters
        hipLaunchKernelGGL(k_split_plan, dim3(1), dim3(kSplitPlanThreads), 0, stream, U.split_cost, tasks,
                           U.split_flag, U.split_list, U.split_count, U.split_sub_done, U.chunk_done, units);
      } else {
        hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((units + 255) / 256)), dim3(256), 0, stream, U.chunk_done,
                           (uint32_t)units, 0u);
        T.split_flag = nullptr;
        if (have_split)  // no steps of these tasks yet: record them, split nothing
          hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((tasks + 255) / 256)), dim3(256), 0, stream, U.split_cost,
                             (uint32_t)tasks, 0u);
        else
          T.split_cost = nullptr;
      }
      const uint64_t waves = tasks + (T.split_flag ? (uint64_t)kSplitSubWaves : 0u);
      const dim3 grid((unsigned)((waves + (kTraceThreads / 64) - 1) / (kTraceThreads / 64))), block(kTraceThreads);
      const bool scalar = U.traversal == kUnitPacketScalar;
      // the variant without the split's code keeps the registers of the plain task kernel
      auto kernel = U.samples ? (scalar ? k_render_tasks<true, true, false> : k_render_tasks<true, false, false>)
                              : (scalar ? k_render_tasks<false, true, false> : k_render_tasks<false, false, false>);
      if (T.split_cost)
        kernel = U.samples ? (scalar ? k_render_tasks<true, true, true> : k_render_tasks<true, false, true>)
                           : (scalar ? k_render_tasks<false, true, true> : k_render_tasks<false, false, true>);
      hipLaunchKernelGGL(kernel, grid, block, 0, stream, v, T, (int)chunks);
      g_last_kernel = "k_render_tasks";
      return check_launch("render_tasks", err);
    }
    // per-ray traversal: the chunk kernel, one wave per chunk, an LDS stack
    // column per lane
    const size_t chunk_lds = stack_bytes(scene);
    if (chunk_lds > 160 * 1024) {
      if (err) *err = "render_chunks: BVH too deep for the LDS traversal stack";
      return false;
    }
    if (tasks / (kTraceThreads / 64) >= 0x7fffffffull) {
      if (err) *err = "render_chunks: too many chunks for one launch";
      return false;
    }
    const int chunk_blocks = (int)((tasks + (kTraceThreads / 64) - 1) / (kTraceThreads / 64));
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((units + 255) / 256)), dim3(256), 0, stream, U.chunk_done,
                       (uint32_t)units, 0u);
    launch_chunks(U, v, chunk_blocks, chunk_lds, (int)chunks, stream);
    g_last_kernel = "k_render_chunks";
    return check_launch("render_chunks", err);
  }
  // units per queue grab: as given, or (0) as many as keep two grabs per
  // resident wave, at most auto_grab_max
  UnitLaunch G = U;
  if (G.units_per_grab <= 0)
    G.units_per_grab = (int)std::min<uint64_t>((uint64_t)std::max(G.auto_grab_max, 1),
                                               std::max<uint64_t>(1, units / (2 * resident_waves)));
  G.units_per_grab = std::min(G.units_per_grab, kMaxUnitsPerGrab);
  const uint64_t need = (units + (uint64_t)G.units_per_grab - 1) / (uint64_t)G.units_per_grab;
  const uint64_t waves = (uint64_t)blocks * (kTraceThreads / 64);
  if (waves > need) blocks = (int)((need + (kTraceThreads / 64) - 1) / (kTraceThreads / 64));
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, stream, G.queue, (uint32_t)kQueueWords, 0u);
  if (G.traversal == kUnitPerRay)
    launch_units_variant<kUnitPerRay>(G, v, blocks, lds, stream);
  else if (G.traversal == kUnitPacketScalar)
    launch_units_variant<kUnitPacketScalar>(G, v, blocks, lds, stream);
  else
    launch_units_variant<kUnitPacket>(G, v, blocks, lds, stream);
  g_last_kernel = "k_render_units";
  return check_launch("render_units", err);
}

const char* last_render_kernel() { return g_last_kernel; }
void set_split_launches(bool on) { g_split = on; }

void set_trace_mode(int mode) { g_trace_mode = mode; }
int trace_mode() { return g_trace_mode; }
int unit_traversal_of(int mode) {
  return mode == kTracePerRay ? kUnitPerRay : mode == kTracePacketScalar ? kUnitPacketScalar : kUnitPacket;
}
int trace_mode_of(int traversal) {
  return traversal == kUnitPerRay ? kTracePerRay : traversal == kUnitPacketScalar ? kTracePacketScalar : kTracePacket;
}

bool launch_copy_probe(const float4* src, float4* dst, size_t n, hipStream_t stream, std::string* err) {
  if (n == 0) return true;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 256 * 32);
  hipLaunchKernelGGL(k_copy_probe, dim3((unsigned)blocks), dim3(256), 0, stream, src, dst, n);
  return check_launch("copy_probe", err);
}

bool launch_trace_stats(const DeviceScene& scene, const EyeLaunch& L, unsigned long long* totals,
                        hipStream_t stream, std::string* err) {
  const uint64_t total = L.rays_per_pose * (uint64_t)L.pose_count;
  if (total == 0) return true;
  const size_t lds = stack_bytes(scene);
  if (lds > 160 * 1024) {
    if (err) *err = "trace: BVH too deep for the LDS traversal stack";
    return false;
  }
  const uint64_t blocks = (total + kTraceThreads - 1) / kTraceThreads;
  hipLaunchKernelGGL(k_trace_stats, dim3((unsigned)blocks), dim3(kTraceThreads), lds, stream, view_of(scene), L,
                     total, totals);
  return check_launch("trace_stats", err);
}

bool launch_frame_single_dim_fast(const float4* colours, int omm_count, uchar4* frame, int width,
                                  hipStream_t stream, std::string* err) {
  if (width <= 0 || omm_count <= 0) return true;
  hipLaunchKernelGGL(k_frame_sdf, dim3((width + 255) / 256), dim3(256), 0, stream, colours, omm_count, frame, width);
  return check_launch("frame", err);
}

bool launch_frame_single_dim(const float4* colours, int omm_count, uchar4* frame, int width, int height,
                             hipStream_t stream, std::string* err) {
  if (width <= 0 || height <= 0 || omm_count <= 0) return true;
  hipLaunchKernelGGL(k_frame_sd, dim3((width + 255) / 256, height), dim3(256), 0, stream, colours, omm_count,
                     frame, width, height);
  return check_launch("frame", err);
}

bool launch_frame_raw_samples(const float4* samples, const uint32_t* offsets, int omm_count, int spp,
                              uchar4* frame, int width, int height, hipStream_t stream, std::string* err) {
  if (width <= 0 || height <= 0 || omm_count <= 0) return true;
  hipLaunchKernelGGL(k_frame_raw, dim3((width + 255) / 256, height), dim3(256), 0, stream, samples, offsets,
                     omm_count, spp, frame, width, height);
  return check_launch("frame", err);
}

bool launch_camera_frame(const DeviceScene& scene, const CameraLaunch& C, uchar4* frame, Ray* rays,
                         hipStream_t stream, std::string* err) {
  if (C.width <= 0 || C.height <= 0) return true;
  const long long pixels = (long long)C.width * C.height;
  if (pixels > 0x7fffffffLL) {
    if (err) *err = "camera: frame too large";
    return false;
  }
  const size_t lds = stack_bytes(scene);
  if (lds > 160 * 1024) {
    if (err) *err = "camera: BVH too deep for the LDS traversal stack";
    return false;
  }
  hipLaunchKernelGGL(k_camera_frame, dim3((unsigned)((pixels + kTraceThreads - 1) / kTraceThreads)),
                     dim3(kTraceThreads), lds, stream, view_of(scene), C, frame, rays);
  return check_launch("camera_frame", err);
}

bool launch_generate_rays(const EyeLaunch& L, Ray* rays, hipStream_t stream, std::string* err) {
  const uint64_t total = L.rays_per_pose * (uint64_t)L.pose_count;
  if (total == 0) return true;
  hipLaunchKernelGGL(k_generate_rays, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, L, total, rays);
  return check_launch("generate_rays", err);
}

bool launch_packet_guard_probe(const DeviceScene& scene, const Ray* rays, int active_lanes, int32_t* ids,
                               hipStream_t stream, std::string* err) {
  if (scene.stack_bound > kPacketStackMax) {
    if (err) *err = "packet probe: BVH too deep for the packet traversal";
    return false;
  }
  hipLaunchKernelGGL(k_packet_guard_probe, dim3(1), dim3(64), 0, stream, view_of(scene), rays, active_lanes, ids);
  return check_launch("packet_guard_probe", err);
}

bool launch_trace_rays(const DeviceScene& scene, const Ray* rays, uint64_t n, int32_t* hit, float4* tbb,
                       hipStream_t stream, std::string* err) {
  if (n == 0) return true;
  const size_t lds = stack_bytes(scene);
  if (lds > 160 * 1024) {
    if (err) *err = "trace: BVH too deep for the LDS traversal stack";
    return false;
  }
  hipLaunchKernelGGL(k_trace_rays, dim3((unsigned)((n + kTraceThreads - 1) / kTraceThreads)), dim3(kTraceThreads),
                     lds, stream, view_of(scene), rays, n, hit, tbb);
  return check_launch("trace_rays", err);
}

}  // namespace eye
