// gfx950 render kernels: fused Gaussian-cone ray generation + 4-wide BVH
// traversal + shading (replacing __raygen__ommatidium, the OptiX traversal,
// __closesthit__radiance and the miss programs, libEyeRenderer3/shaders.cu:664-811),
// and the per-ommatidium wave reduction + projection (shaders.cu:341-406).
#include <hip/hip_runtime.h>

#include <string>

#include "render.h"

namespace eye {
namespace {

constexpr int kTraceThreads = 256;

struct SceneView {
  const WideNode* nodes;
  const float4* tris;
  const float4* cols;
  int node_count;
  int sky_mode;
};

SceneView view_of(const DeviceScene& s) {
  SceneView v;
  v.nodes = s.nodes;
  v.tris = s.tris;
  v.cols = s.cols;
  v.node_count = s.node_count;
  v.sky_mode = s.sky_mode;
  return v;
}

struct Hit {
  float t, b1, b2;
  uint32_t k;   // sorted triangle position
  uint32_t id;  // original triangle id, 0xffffffff for a miss
};

__device__ __forceinline__ void test_leaf(const float4* __restrict__ tris, int32_t code, const RayPre& rp,
                                          Hit& h) {
  const uint32_t first = leaf_first(code);
  const uint32_t cnt = leaf_count(code);
  for (uint32_t q = 0; q < cnt; ++q) {
    const uint32_t k = first + q;
    const float4 a = tris[3 * (size_t)k + 0];
    const float4 b = tris[3 * (size_t)k + 1];
    const float4 c = tris[3 * (size_t)k + 2];
    float t, b1, b2;
    if (intersect_triangle(rp, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), &t, &b1, &b2)) {
      const uint32_t id = __float_as_uint(a.w);
      if (t < h.t || (t == h.t && id < h.id)) {
        h.t = t;
        h.b1 = b1;
        h.b2 = b2;
        h.k = k;
        h.id = id;
      }
    }
  }
}

// Closest hit over the 4-wide BVH.  `stack` points at this lane's column of
// an LDS array laid out [depth][blockDim.x].  The result is independent of
// the traversal order: the (t, id) pair is minimised lexicographically.
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
    const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
    const int4 ch = reinterpret_cast<const int4*>(np)[6];
    float te[4];
    const float limit = h.t * kCullSlack;
    te[0] = slab_enter<kParallel>(rp, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, limit);
    te[1] = slab_enter<kParallel>(rp, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, limit);
    te[2] = slab_enter<kParallel>(rp, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, limit);
    te[3] = slab_enter<kParallel>(rp, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, limit);
    const int32_t cc[4] = {ch.x, ch.y, ch.z, ch.w};
    int ni = 0;
    int ic[4];
    float it[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int32_t c = cc[s];
      if (c == kChildEmpty || !(te[s] < INFINITY)) continue;
      if (child_is_leaf(c)) {
        if (kStats) *tris_tested += leaf_count(c);
        test_leaf(S.tris, c, rp, h);
      } else {
        ic[ni] = c;
        it[ni] = te[s];
        ++ni;
      }
    }
    if (ni == 0) {
      if (sp == 0) break;
      --sp;
      node = (int)stack[sp * stride];
      continue;
    }
    // order internal children near to far (insertion sort, at most 4)
    for (int a = 1; a < ni; ++a) {
      const float tv = it[a];
      const int cv = ic[a];
      int b = a - 1;
      while (b >= 0 && it[b] > tv) {
        it[b + 1] = it[b];
        ic[b + 1] = ic[b];
        --b;
      }
      it[b + 1] = tv;
      ic[b + 1] = cv;
    }
    for (int a = ni - 1; a >= 1; --a) {
      stack[sp * stride] = (uint32_t)ic[a];
      ++sp;
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
}

__device__ __forceinline__ V3 shade(const SceneView& S, const Hit& h, float dx, float dy, float dz) {
  if (h.id == 0xffffffffu) return sky_colour(S.sky_mode, dx, dy, dz);
  const float4 c0 = S.cols[3 * (size_t)h.k + 0];
  const float4 c1 = S.cols[3 * (size_t)h.k + 1];
  const float4 c2 = S.cols[3 * (size_t)h.k + 2];
  return shade_hit(h.b1, h.b2, v3(c0.x, c0.y, c0.z), v3(c1.x, c1.y, c1.z), v3(c2.x, c2.y, c2.z),
                   __float_as_uint(c0.w));
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

__global__ __launch_bounds__(kTraceThreads) void k_trace_samples(SceneView S, EyeLaunch L, uint64_t total,
                                                                 float4* __restrict__ samples) {
  extern __shared__ uint32_t lds_stack[];
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= total) return;
  int p, i, s, n;
  locate_ray(L, r, &p, &i, &s, &n);
  const Ommatidium o = L.omms[i];
  const Pose pose = L.poses[p];
  const U4 rnd = sample_randoms(L.seed, L.eye, (uint32_t)i, (uint32_t)s, L.frame, L.pose_base + (uint32_t)p);
  const Ray ray = generate_ray(o, pose, rnd);
  const RayPre rp = precompute_ray(ray);
  Hit h;
  closest_hit(S, rp, lds_stack + threadIdx.x, blockDim.x, h);
  const V3 c = shade(S, h, ray.dx, ray.dy, ray.dz);
  samples[r] = make_float4(c.x, c.y, c.z, h.id == 0xffffffffu ? -1.0f : h.t);
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

// One wave per (pose, ommatidium).
__global__ __launch_bounds__(256) void k_reduce(EyeLaunch L, const float4* __restrict__ samples,
                                                float4* __restrict__ colours) {
  const int lane = threadIdx.x & 63;
  const uint64_t unit = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint64_t units = (uint64_t)L.omm_count * (uint64_t)L.pose_count;
  if (unit >= units) return;
  const uint64_t p = unit / (uint64_t)L.omm_count;
  const int i = (int)(unit - p * (uint64_t)L.omm_count);
  uint32_t first, n;
  if (L.offsets == nullptr) {
    first = (uint32_t)i * (uint32_t)L.spp;
    n = (uint32_t)L.spp;
  } else {
    first = L.offsets[i];
    n = L.offsets[i + 1] - first;
  }
  const float inv = 1.0f / (float)(n > 0 ? n : 1u);
  const float4* base = samples + p * L.rays_per_pose + first;
  float r = 0.0f, g = 0.0f, b = 0.0f;
  for (uint32_t s = (uint32_t)lane; s < n; s += 64) {
    const float4 v = base[s];
    r = r + v.x * inv;
    g = g + v.y * inv;
    b = b + v.z * inv;
  }
  for (int m = 32; m >= 1; m >>= 1) {
    r = r + __shfl_xor(r, m);
    g = g + __shfl_xor(g, m);
    b = b + __shfl_xor(b, m);
  }
  if (lane == 0) colours[unit] = make_float4(r, g, b, (float)n);
}

__device__ __forceinline__ uchar4 make_color(float4 c) {
  return make_uchar4(encode_channel(c.x), encode_channel(c.y), encode_channel(c.z), 255);
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

__global__ void k_frame_raw(const float4* __restrict__ samples, int omm_count, int spp,
                            uchar4* __restrict__ frame, int width, int height) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= width || y >= height || x >= omm_count || y >= spp) return;
  // the reference stores c / spp per sample and displays it unsummed
  const float4 v = samples[(size_t)x * spp + y];
  const float inv = 1.0f / (float)spp;
  frame[(size_t)y * width + x] = make_color(make_float4(v.x * inv, v.y * inv, v.z * inv, 0.0f));
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

bool launch_trace_samples(const DeviceScene& scene, const EyeLaunch& L, float4* samples, hipStream_t stream,
                          std::string* err) {
  const uint64_t total = L.rays_per_pose * (uint64_t)L.pose_count;
  if (total == 0) return true;
  const size_t lds = stack_bytes(scene);
  if (lds > 160 * 1024) {
    if (err) *err = "trace: BVH too deep for the LDS traversal stack";
    return false;
  }
  const uint64_t blocks = (total + kTraceThreads - 1) / kTraceThreads;
  hipLaunchKernelGGL(k_trace_samples, dim3((unsigned)blocks), dim3(kTraceThreads), lds, stream, view_of(scene), L,
                     total, samples);
  return check_launch("trace_samples", err);
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

bool launch_reduce(const EyeLaunch& L, const float4* samples, float4* colours, hipStream_t stream,
                   std::string* err) {
  const uint64_t units = (uint64_t)L.omm_count * (uint64_t)L.pose_count;
  if (units == 0) return true;
  const uint64_t blocks = (units + 3) / 4;
  hipLaunchKernelGGL(k_reduce, dim3((unsigned)blocks), dim3(256), 0, stream, L, samples, colours);
  return check_launch("reduce", err);
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

bool launch_frame_raw_samples(const float4* samples, int omm_count, int spp, uchar4* frame, int width,
                              int height, hipStream_t stream, std::string* err) {
  if (width <= 0 || height <= 0 || omm_count <= 0) return true;
  hipLaunchKernelGGL(k_frame_raw, dim3((width + 255) / 256, height), dim3(256), 0, stream, samples, omm_count, spp,
                     frame, width, height);
  return check_launch("frame", err);
}

bool launch_generate_rays(const EyeLaunch& L, Ray* rays, hipStream_t stream, std::string* err) {
  const uint64_t total = L.rays_per_pose * (uint64_t)L.pose_count;
  if (total == 0) return true;
  hipLaunchKernelGGL(k_generate_rays, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, L, total, rays);
  return check_launch("generate_rays", err);
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
