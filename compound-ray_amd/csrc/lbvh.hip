// On-GPU LBVH build for gfx950: Morton codes, LDS-staged stable radix sort,
// Karras hierarchy, bottom-up refit, breadth-first collapse to a kBvhWidth-wide BVH.
// The rules are the ones documented in bvh_format.h; this file is the
// parallel implementation.  Replaces the OptiX GAS/IAS build of
// libEyeRenderer3/MulticamScene.cpp:1052-1428.
#include <hip/hip_runtime.h>

#include <chrono>
#include <string>
#include <vector>

#include "device_scene.h"
#include "scan.h"

namespace eye {
namespace {

#define LB_TRY(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      if (err) *err = std::string("lbvh: ") + #x + ": " + hipGetErrorString(e_);   \
      return false;                                                                \
    }                                                                              \
  } while (0)

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kRadix = 256;
constexpr int kWaves = kSortThreads / 64;

// Deterministic min/max (ties return the second operand), identical on the CPU.
__device__ __forceinline__ float dmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float dmax(float a, float b) { return a > b ? a : b; }

// Order-preserving map of floats to unsigned keys (-0 sorts below +0).
__device__ __forceinline__ uint32_t fkey(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ inline float fkey_inv(uint32_t k) {
  const uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  float f;
  memcpy(&f, &b, 4);
  return f;
}

// Triangle boxes and centroids; each block also writes the order-preserving
// keys of its centroid box to part[6 * blockIdx.x ...] (lo x,y,z, hi x,y,z),
// reduced by k_centroid_box.  No atomics: the six global addresses would
// otherwise serialise one atomic per wave.
__global__ __launch_bounds__(256) void k_tri_bounds(const float* __restrict__ pos, int n, float* __restrict__ tb,
                                                    float* __restrict__ cen, uint32_t* __restrict__ part) {
  __shared__ uint32_t wave_keys[4][6];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t klo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu};
  uint32_t khi[3] = {0u, 0u, 0u};
  if (i < n) {
    const float* p = pos + 9 * (size_t)i;
    for (int a = 0; a < 3; ++a) {
      const float lo = dmin(dmin(p[a], p[3 + a]), p[6 + a]);
      const float hi = dmax(dmax(p[a], p[3 + a]), p[6 + a]);
      tb[6 * (size_t)i + a] = lo;
      tb[6 * (size_t)i + 3 + a] = hi;
      const float c = (lo + hi) * 0.5f;
      cen[3 * (size_t)i + a] = c;
      klo[a] = fkey(c);
      khi[a] = fkey(c);
    }
  }
  const int wave = threadIdx.x >> 6;
  for (int a = 0; a < 3; ++a) {
    uint32_t lo = klo[a], hi = khi[a];
    for (int m = 32; m >= 1; m >>= 1) {
      const uint32_t olo = __shfl_xor(lo, m);
      const uint32_t ohi = __shfl_xor(hi, m);
      lo = olo < lo ? olo : lo;
      hi = ohi > hi ? ohi : hi;
    }
    if ((threadIdx.x & 63) == 0) {
      wave_keys[wave][a] = lo;
      wave_keys[wave][3 + a] = hi;
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    uint32_t k = wave_keys[0][a];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      const uint32_t o = wave_keys[w][a];
      k = a < 3 ? (o < k ? o : k) : (o > k ? o : k);
    }
    part[6 * (size_t)blockIdx.x + a] = k;
  }
}

// One block folds the per-block centroid keys into cbounds[6].
__global__ __launch_bounds__(256) void k_centroid_box(const uint32_t* __restrict__ part, int parts,
                                                      uint32_t* __restrict__ cbounds) {
  __shared__ uint32_t wave_keys[4][6];
  uint32_t k[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  for (int j = threadIdx.x; j < parts; j += blockDim.x)
    for (int a = 0; a < 6; ++a) {
      const uint32_t o = part[6 * (size_t)j + a];
      k[a] = a < 3 ? (o < k[a] ? o : k[a]) : (o > k[a] ? o : k[a]);
    }
  for (int a = 0; a < 6; ++a)
    for (int m = 32; m >= 1; m >>= 1) {
      const uint32_t o = __shfl_xor(k[a], m);
      k[a] = a < 3 ? (o < k[a] ? o : k[a]) : (o > k[a] ? o : k[a]);
    }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 6; ++a) wave_keys[wave][a] = k[a];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    uint32_t r = wave_keys[0][a];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      const uint32_t o = wave_keys[w][a];
      r = a < 3 ? (o < r ? o : r) : (o > r ? o : r);
    }
    cbounds[a] = r;
  }
}

__global__ void k_morton(const float* __restrict__ cen, int n, const uint32_t* __restrict__ cbounds,
                         uint32_t* __restrict__ codes, uint32_t* __restrict__ ids) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = fkey_inv(cbounds[a]);
    hi[a] = fkey_inv(cbounds[3 + a]);
  }
  codes[i] = morton30(cen[3 * (size_t)i], cen[3 * (size_t)i + 1], cen[3 * (size_t)i + 2], lo, hi);
  ids[i] = (uint32_t)i;
}

// Per-tile digit histogram, written digit-major: hist[digit * tiles + tile].
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint32_t* __restrict__ keys, int n,
                                                            int shift, uint32_t* __restrict__ hist,
                                                            int tiles) {
  __shared__ uint32_t h[kRadix];
  h[threadIdx.x] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kSortTile;
  for (int j = 0; j < kSortItems; ++j) {
    const size_t idx = base + (size_t)j * kSortThreads + threadIdx.x;
    if (idx < (size_t)n) atomicAdd(&h[(keys[idx] >> shift) & (kRadix - 1)], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter.  Items are taken in the order (item j, wave, lane); ranks
// among equal digits inside a wave come from ballot matching, across waves
// and items from LDS running counters, across tiles from the scanned
// histogram.
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int n, int shift, const uint32_t* __restrict__ offs, int tiles) {
  __shared__ uint32_t wave_cnt[kWaves][kRadix];
  __shared__ uint32_t wave_off[kWaves][kRadix];
  __shared__ uint32_t run_base[kRadix];
  __shared__ uint32_t g_off[kRadix];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  run_base[tid] = 0;
  g_off[tid] = offs[(size_t)tid * tiles + blockIdx.x];
  const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const size_t base = (size_t)blockIdx.x * kSortTile;
  for (int j = 0; j < kSortItems; ++j) {
    const size_t idx = base + (size_t)j * kSortThreads + tid;
    const bool valid = idx < (size_t)n;
    const uint32_t key = valid ? kin[idx] : 0u;
    const uint32_t val = valid ? vin[idx] : 0u;
    const uint32_t d = (key >> shift) & (kRadix - 1);
    for (int k = 0; k < kWaves; ++k) wave_cnt[k][tid] = 0;
    __syncthreads();
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    uint32_t rank = 0;
    if (valid) {
      rank = (uint32_t)__popcll(peers & lt_mask);
      const int leader = 63 - __clzll((long long)peers);
      if (leader == lane) wave_cnt[w][d] = (uint32_t)__popcll(peers);
    }
    __syncthreads();
    {
      uint32_t run = run_base[tid];
      for (int k = 0; k < kWaves; ++k) {
        const uint32_t c = wave_cnt[k][tid];
        wave_off[k][tid] = run;
        run += c;
      }
      run_base[tid] = run;
    }
    __syncthreads();
    if (valid) {
      const uint32_t pos = g_off[d] + wave_off[w][d] + rank;
      if (pos < (uint32_t)n) {  // always true for a consistent histogram; never write out of bounds
        kout[pos] = key;
        vout[pos] = val;
      }
    }
  }
}

__global__ void k_karras(const uint32_t* __restrict__ codes, int n, int32_t* __restrict__ left,
                         int32_t* __restrict__ right, int32_t* __restrict__ parent,
                         int32_t* __restrict__ range) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {  // leaf ranges
    range[2 * (size_t)(n - 1 + i)] = i;
    range[2 * (size_t)(n - 1 + i) + 1] = i;
  }
  if (i >= n - 1) return;
  int l, r, first, last;
  bool ll, rl;
  karras_node(codes, n, i, &l, &ll, &r, &rl, &first, &last);
  const int lc = ll ? (n - 1 + l) : l;
  const int rc = rl ? (n - 1 + r) : r;
  left[i] = lc;
  right[i] = rc;
  parent[lc] = i;
  parent[rc] = i;
  range[2 * (size_t)i] = first;
  range[2 * (size_t)i + 1] = last;
  if (i == 0) parent[0] = -1;
}

// Bottom-up refit without inter-wave signalling.  Release/acquire fences at
// device scope write back and invalidate the XCD-local L2 on this part, so
// the refit orders itself with launch boundaries instead: internal node
// depths come from pointer jumping over the parent array, then one launch
// per depth, deepest first, unions the two child boxes of that depth's nodes.
__global__ void k_leaf_bounds(const uint32_t* __restrict__ ids, const float* __restrict__ tb, int n,
                              float* __restrict__ bounds) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const size_t leaf = (size_t)(n - 1 + k);
  const uint32_t id = ids[k];
  for (int a = 0; a < 6; ++a) bounds[6 * leaf + a] = tb[6 * (size_t)id + a];
}

// up[i] = parent of internal node i (-1 at the root), dist[i] = 1 (0 at the root).
__global__ void k_depth_init(const int32_t* __restrict__ parent, int m, int32_t* __restrict__ up,
                             int32_t* __restrict__ dist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int32_t p = parent[i];
  const bool valid = p >= 0 && p < m;
  up[i] = valid ? p : -1;
  dist[i] = valid ? 1 : 0;
}

// One pointer-jumping step: afterwards up[i] is twice as far above i, and
// dist[i] counts the edges from i to it.  Once up[i] is -1, dist[i] is the depth.
__global__ void k_depth_jump(const int32_t* __restrict__ up_in, const int32_t* __restrict__ dist_in, int m,
                             int32_t* __restrict__ up_out, int32_t* __restrict__ dist_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int32_t u = up_in[i];
  if (u < 0) {
    up_out[i] = -1;
    dist_out[i] = dist_in[i];
  } else {
    up_out[i] = up_in[u];
    dist_out[i] = dist_in[i] + dist_in[u];
  }
}

__global__ void k_refit_level(const int32_t* __restrict__ depth, int level, int m, const int32_t* __restrict__ left,
                              const int32_t* __restrict__ right, float* __restrict__ bounds) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || depth[i] != level) return;
  const size_t l = (size_t)left[i], r = (size_t)right[i];
  float b[6];
  for (int a = 0; a < 3; ++a) {
    b[a] = dmin(bounds[6 * l + a], bounds[6 * r + a]);
    b[3 + a] = dmax(bounds[6 * l + 3 + a], bounds[6 * r + 3 + a]);
  }
  for (int a = 0; a < 6; ++a) bounds[6 * (size_t)i + a] = b[a];
}

__global__ void k_collapse_count(const int32_t* __restrict__ srcs, int count, BinaryTreeView t,
                                 int leaf_max, uint32_t* __restrict__ internal_count) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  int ch[kBvhWidth];
  const int c = collapse_children(t, srcs[j], leaf_max, ch);
  uint32_t k = 0;
  for (int s = 0; s < c; ++s) k += wide_child_is_internal(t, ch[s], leaf_max) ? 1u : 0u;
  internal_count[j] = k;
}

__global__ void k_collapse_emit(const int32_t* __restrict__ srcs, int count, BinaryTreeView t,
                                int leaf_max, const uint32_t* __restrict__ offs, int base,
                                int next_base, int32_t* __restrict__ next_srcs,
                                WideNode* __restrict__ nodes, int32_t* __restrict__ depth,
                                int32_t* __restrict__ stack_bound) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  int ch[kBvhWidth];
  const int c = collapse_children(t, srcs[j], leaf_max, ch);
  const int w = base + j;
  const int dw = depth[w];
  int k_internal = 0;
  for (int s = 0; s < c; ++s) k_internal += wide_child_is_internal(t, ch[s], leaf_max) ? 1 : 0;
  WideNode node;
  int next = 0;
  for (int s = 0; s < kBvhWidth; ++s) {
    WideHalf& hf = node.half[s / kHalfWidth];
    const int q = s % kHalfWidth;
    if (s < c) {
      const int b = ch[s];
      const float* bb = t.bounds + 6 * (size_t)b;
      hf.lo_x[q] = bb[0]; hf.lo_y[q] = bb[1]; hf.lo_z[q] = bb[2];
      hf.hi_x[q] = bb[3]; hf.hi_y[q] = bb[4]; hf.hi_z[q] = bb[5];
      if (wide_child_is_internal(t, b, leaf_max)) {
        const int slot = (int)offs[j] + next;
        ++next;
        hf.child[q] = next_base + slot;
        next_srcs[slot] = b;
        depth[next_base + slot] = dw + (k_internal - 1);
      } else {
        hf.child[q] = encode_leaf((uint32_t)t.range[2 * (size_t)b], (uint32_t)bin_size(t, b));
      }
    } else {
      hf.lo_x[q] = hf.hi_x[q] = kEmptyBox;
      hf.lo_y[q] = hf.hi_y[q] = kEmptyBox;
      hf.lo_z[q] = hf.hi_z[q] = kEmptyBox;
      hf.child[q] = kChildEmpty;
    }
    hf.pad[q] = 0;
  }
  nodes[w] = node;
  atomicMax(stack_bound, dw + (k_internal > 1 ? k_internal - 1 : 0));
}

__global__ void k_gather(const float* __restrict__ pos, const float* __restrict__ col,
                         const uint32_t* __restrict__ mode, const uint32_t* __restrict__ ids, int n,
                         float4* __restrict__ tris, float4* __restrict__ cols) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t id = ids[k];
  const float* p = pos + 9 * (size_t)id;
  const float* c = col + 9 * (size_t)id;
  tris[3 * (size_t)k + 0] = make_float4(p[0], p[1], p[2], __uint_as_float(id));
  tris[3 * (size_t)k + 1] = make_float4(p[3], p[4], p[5], 0.0f);
  tris[3 * (size_t)k + 2] = make_float4(p[6], p[7], p[8], 0.0f);
  cols[3 * (size_t)k + 0] = make_float4(c[0], c[1], c[2], __uint_as_float(mode[id]));
  cols[3 * (size_t)k + 1] = make_float4(c[3], c[4], c[5], 0.0f);
  cols[3 * (size_t)k + 2] = make_float4(c[6], c[7], c[8], 0.0f);
}

inline int blocks_for(size_t n, int threads) { return (int)((n + threads - 1) / threads); }

// Host check of the finished structure before anything indexes through it:
// the sort order is a permutation, every internal child index points forward
// to a node referenced exactly once, and the leaves cover every triangle
// exactly once.  A violation is reported as an error instead of letting a
// kernel read out of bounds.
// Structure check of a finished build, on the device: the sorted order is a
// permutation, every leaf range lies inside the triangle array, every
// triangle is in exactly one leaf, every internal child index points forward
// and every node but the root has exactly one parent.  Traversal relies on
// all of it to stay in bounds.  Violations set bits of `flags`.
enum : uint32_t { kBadOrder = 1u, kBadLeaf = 2u, kTwoLeaves = 4u, kNoLeaf = 8u, kBadChild = 16u, kUnreached = 32u };

__global__ void k_count_order(const uint32_t* __restrict__ ids, int n, uint32_t* __restrict__ seen,
                              uint32_t* __restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t v = ids[i];
  if (v >= (uint32_t)n)
    atomicOr(flags, kBadOrder);
  else
    atomicAdd(&seen[v], 1u);
}

__global__ void k_count_refs(const WideNode* __restrict__ nodes, int node_count, int n, uint32_t* __restrict__ refs,
                             uint32_t* __restrict__ covered, uint32_t* __restrict__ flags) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= node_count) return;
  for (int s = 0; s < kBvhWidth; ++s) {
    const int32_t c = nodes[w].half[s / kHalfWidth].child[s % kHalfWidth];
    if (c == kChildEmpty) continue;
    if (child_is_leaf(c)) {
      const uint32_t f = leaf_first(c), k = leaf_count(c);
      if ((uint64_t)f + k > (uint64_t)n) {
        atomicOr(flags, kBadLeaf);
        continue;
      }
      for (uint32_t q = f; q < f + k; ++q) atomicAdd(&covered[q], 1u);
    } else if (c <= w || c >= node_count) {
      atomicOr(flags, kBadChild);
    } else {
      atomicAdd(&refs[c], 1u);
    }
  }
}

__global__ void k_check_counts(const uint32_t* __restrict__ seen, const uint32_t* __restrict__ covered, int n,
                               const uint32_t* __restrict__ refs, int node_count, uint32_t* __restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t bad = 0;
  if (i < n) {
    if (seen[i] != 1u) bad |= kBadOrder;
    if (covered[i] == 0u) bad |= kNoLeaf;
    if (covered[i] > 1u) bad |= kTwoLeaves;
  }
  if (i >= 1 && i < node_count) {
    if (refs[i] == 0u) bad |= kUnreached;
    if (refs[i] > 1u) bad |= kBadChild;
  }
  if (bad) atomicOr(flags, bad);
}

std::string structure_error(uint32_t bits) {
  if (bits & kBadOrder) return "lbvh: sorted order is not a permutation";
  if (bits & kBadLeaf) return "lbvh: leaf range out of bounds";
  if (bits & kTwoLeaves) return "lbvh: triangle referenced by two leaves";
  if (bits & kNoLeaf) return "lbvh: triangle not covered by any leaf";
  if (bits & kBadChild) return "lbvh: bad internal child index";
  if (bits & kUnreached) return "lbvh: unreachable wide node";
  return "";
}

}  // namespace

void free_device_scene(DeviceScene* s) {
  if (s->nodes) (void)hipFree(s->nodes);
  if (s->tris) (void)hipFree(s->tris);
  if (s->cols) (void)hipFree(s->cols);
  if (s->sorted_ids) (void)hipFree(s->sorted_ids);
  if (s->codes) (void)hipFree(s->codes);
  *s = DeviceScene();
}

bool build_device_scene(const float* d_pos, const float* d_col, const uint32_t* d_mode, int n,
                        int leaf_max, hipStream_t stream, DeviceScene* out, std::string* err) {
  const auto t0 = std::chrono::steady_clock::now();
  free_device_scene(out);
  out->leaf_max = leaf_max;
  out->tri_count = n;
  if (leaf_max < 1 || leaf_max > kLeafMaxLimit) {
    if (err) *err = "lbvh: leaf_max must be in [1, 16]";
    return false;
  }
  if (n < 0 || n >= (1 << 27)) {
    if (err) *err = "lbvh: triangle count must be below 2^27";
    return false;
  }
  if (n == 0) {
    out->node_count = 0;
    return true;
  }
  const int nb = 2 * n - 1;  // binary nodes, combined indexing
  const int tiles = blocks_for(n, kSortTile);
  float* tb = nullptr;
  float* cen = nullptr;
  uint32_t* cbounds = nullptr;
  uint32_t *codes_a = nullptr, *codes_b = nullptr, *ids_a = nullptr, *ids_b = nullptr;
  uint32_t* hist = nullptr;
  int32_t *left = nullptr, *right = nullptr, *parent = nullptr, *range = nullptr;
  float* bounds = nullptr;
  uint32_t* cpart = nullptr;
  int32_t* dist = nullptr;
  uint32_t* vflags = nullptr;
  int32_t *srcs_a = nullptr, *srcs_b = nullptr, *depth = nullptr, *stack_bound = nullptr;
  uint32_t *icount = nullptr, *ioffs = nullptr;
  ScanScratch scan;
  bool ok = false;
  const int max_nodes = n > 1 ? n - 1 : 1;
  int node_total = 0, level_count = 0, host_sb = 0;
  do {
#define LB_STEP(x)                                                                         \
  {                                                                                        \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      if (err) *err = std::string("lbvh: ") + #x + ": " + hipGetErrorString(e_);           \
      break;                                                                               \
    }                                                                                      \
  }
    bool fail = true;
    do {
      LB_STEP(hipMalloc(&tb, sizeof(float) * 6 * (size_t)n));
      LB_STEP(hipMalloc(&cen, sizeof(float) * 3 * (size_t)n));
      LB_STEP(hipMalloc(&cbounds, sizeof(uint32_t) * 6));
      LB_STEP(hipMalloc(&codes_a, sizeof(uint32_t) * (size_t)n));
      LB_STEP(hipMalloc(&codes_b, sizeof(uint32_t) * (size_t)n));
      LB_STEP(hipMalloc(&ids_a, sizeof(uint32_t) * (size_t)n));
      LB_STEP(hipMalloc(&ids_b, sizeof(uint32_t) * (size_t)n));
      LB_STEP(hipMalloc(&hist, sizeof(uint32_t) * (size_t)kRadix * tiles));
      LB_STEP(hipMalloc(&left, sizeof(int32_t) * (size_t)max_nodes));
      LB_STEP(hipMalloc(&right, sizeof(int32_t) * (size_t)max_nodes));
      LB_STEP(hipMalloc(&parent, sizeof(int32_t) * (size_t)nb));
      LB_STEP(hipMalloc(&range, sizeof(int32_t) * 2 * (size_t)nb));
      LB_STEP(hipMalloc(&bounds, sizeof(float) * 6 * (size_t)nb));
      LB_STEP(hipMalloc(&cpart, sizeof(uint32_t) * 6 * (size_t)blocks_for(n, 256)));
      LB_STEP(hipMalloc(&dist, sizeof(int32_t) * (size_t)max_nodes));
      LB_STEP(hipMalloc(&vflags, sizeof(uint32_t)));
      LB_STEP(hipMalloc(&srcs_a, sizeof(int32_t) * (size_t)max_nodes));
      LB_STEP(hipMalloc(&srcs_b, sizeof(int32_t) * (size_t)max_nodes));
      LB_STEP(hipMalloc(&depth, sizeof(int32_t) * (size_t)max_nodes));
      LB_STEP(hipMalloc(&stack_bound, sizeof(int32_t)));
      LB_STEP(hipMalloc(&icount, sizeof(uint32_t) * (size_t)max_nodes));
      LB_STEP(hipMalloc(&ioffs, sizeof(uint32_t) * ((size_t)max_nodes + 1)));
      LB_STEP(hipMalloc(&out->nodes, sizeof(WideNode) * (size_t)max_nodes));
      LB_STEP(hipMalloc(&out->tris, sizeof(float4) * 3 * (size_t)n));
      LB_STEP(hipMalloc(&out->cols, sizeof(float4) * 3 * (size_t)n));
      LB_STEP(hipMalloc(&out->sorted_ids, sizeof(uint32_t) * (size_t)n));
      LB_STEP(hipMalloc(&out->codes, sizeof(uint32_t) * (size_t)n));
      if (!scan_reserve(&scan, (size_t)kRadix * tiles > (size_t)max_nodes + 1 ? (size_t)kRadix * tiles
                                                                              : (size_t)max_nodes + 1,
                        err))
        break;
      fail = false;
    } while (0);
    if (fail) break;
    fail = true;
    do {
      LB_STEP(hipMemsetAsync(stack_bound, 0, sizeof(int32_t), stream));
      LB_STEP(hipMemsetAsync(depth, 0, sizeof(int32_t), stream));
      hipLaunchKernelGGL(k_tri_bounds, dim3(blocks_for(n, 256)), dim3(256), 0, stream, d_pos, n, tb, cen, cpart);
      hipLaunchKernelGGL(k_centroid_box, dim3(1), dim3(256), 0, stream, cpart, blocks_for(n, 256), cbounds);
      hipLaunchKernelGGL(k_morton, dim3(blocks_for(n, 256)), dim3(256), 0, stream, cen, n, cbounds, codes_a, ids_a);
      // 30-bit keys: passes at bits 0, 8, 16, 24
      uint32_t *kin = codes_a, *kout = codes_b, *vin = ids_a, *vout = ids_b;
      bool sort_ok = true;
      for (int shift = 0; shift < 30; shift += 8) {
        hipLaunchKernelGGL(k_sort_hist, dim3(tiles), dim3(kSortThreads), 0, stream, kin, n, shift, hist, tiles);
        if (!scan_exclusive_u32(&scan, hist, (size_t)kRadix * tiles, stream, err)) { sort_ok = false; break; }
        hipLaunchKernelGGL(k_sort_scatter, dim3(tiles), dim3(kSortThreads), 0, stream, kin, vin, kout, vout, n,
                           shift, hist, tiles);
        uint32_t* t = kin; kin = kout; kout = t;
        t = vin; vin = vout; vout = t;
      }
      if (!sort_ok) break;
      // 4 passes: result back in the *_a buffers
      LB_STEP(hipMemcpyAsync(out->codes, kin, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, stream));
      LB_STEP(hipMemcpyAsync(out->sorted_ids, vin, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, stream));
      hipLaunchKernelGGL(k_karras, dim3(blocks_for(n, 256)), dim3(256), 0, stream, kin, n, left, right, parent, range);
      if (n == 1) LB_STEP(hipMemsetAsync(parent, 0xff, sizeof(int32_t), stream));
      hipLaunchKernelGGL(k_leaf_bounds, dim3(blocks_for(n, 256)), dim3(256), 0, stream, vin, tb, n, bounds);
      if (n > 1) {
        // Split prefixes grow strictly downwards and take at most
        // 30 + ceil(log2 n) values (30-bit codes, then index bits), which
        // bounds the depth of internal nodes; n < 2^27 keeps it below 64, so
        // six pointer-jumping steps settle every depth (seven are run).
        const int m = n - 1;
        int log2n = 0;
        while ((1 << log2n) < n) ++log2n;
        const int max_level = 30 + log2n;
        int32_t *up_a = srcs_a, *up_b = srcs_b, *dist_a = dist, *dist_b = (int32_t*)icount;
        hipLaunchKernelGGL(k_depth_init, dim3(blocks_for(m, 256)), dim3(256), 0, stream, parent, m, up_a, dist_a);
        for (int step = 0; step < 7; ++step) {
          hipLaunchKernelGGL(k_depth_jump, dim3(blocks_for(m, 256)), dim3(256), 0, stream, up_a, dist_a, m, up_b,
                             dist_b);
          int32_t* t = up_a; up_a = up_b; up_b = t;
          t = dist_a; dist_a = dist_b; dist_b = t;
        }
        for (int level = max_level; level >= 0; --level)
          hipLaunchKernelGGL(k_refit_level, dim3(blocks_for(m, 256)), dim3(256), 0, stream, dist_a, level, m, left,
                             right, bounds);
      }
      LB_STEP(hipGetLastError());
      BinaryTreeView tv;
      tv.left = left; tv.right = right; tv.bounds = bounds; tv.range = range; tv.n = n;
      // breadth-first collapse
      int32_t root = 0;
      LB_STEP(hipMemcpyAsync(srcs_a, &root, sizeof(int32_t), hipMemcpyHostToDevice, stream));
      int level_n = 1, base = 0;
      int32_t *cur = srcs_a, *nxt = srcs_b;
      bool lvl_ok = true;
      while (level_n > 0) {
        hipLaunchKernelGGL(k_collapse_count, dim3(blocks_for(level_n, 256)), dim3(256), 0, stream, cur, level_n, tv,
                           leaf_max, icount);
        uint32_t total = 0;
        if (!scan_exclusive_u32_total(&scan, icount, ioffs, (size_t)level_n, stream, &total, err)) {
          lvl_ok = false;
          break;
        }
        const int next_base = base + level_n;
        if ((size_t)next_base + total > (size_t)max_nodes) {
          if (err) *err = "lbvh: wide node count exceeds the binary bound";
          lvl_ok = false;
          break;
        }
        hipLaunchKernelGGL(k_collapse_emit, dim3(blocks_for(level_n, 256)), dim3(256), 0, stream, cur, level_n, tv,
                           leaf_max, ioffs, base, next_base, nxt, out->nodes, depth, stack_bound);
        base = next_base;
        level_n = (int)total;
        int32_t* t = cur; cur = nxt; nxt = t;
        ++level_count;
      }
      if (!lvl_ok) break;
      node_total = base;
      // structure check (seen, covered: the sort keys' buffers, free after the
      // hierarchy; refs: the per-level child counts, free after the collapse)
      LB_STEP(hipMemsetAsync(codes_a, 0, sizeof(uint32_t) * (size_t)n, stream));
      LB_STEP(hipMemsetAsync(codes_b, 0, sizeof(uint32_t) * (size_t)n, stream));
      LB_STEP(hipMemsetAsync(icount, 0, sizeof(uint32_t) * (size_t)max_nodes, stream));
      LB_STEP(hipMemsetAsync(vflags, 0, sizeof(uint32_t), stream));
      hipLaunchKernelGGL(k_count_order, dim3(blocks_for(n, 256)), dim3(256), 0, stream, vin, n, codes_a, vflags);
      hipLaunchKernelGGL(k_count_refs, dim3(blocks_for(node_total, 256)), dim3(256), 0, stream, out->nodes,
                         node_total, n, icount, codes_b, vflags);
      hipLaunchKernelGGL(k_check_counts, dim3(blocks_for(std::max(n, node_total), 256)), dim3(256), 0, stream,
                         codes_a, codes_b, n, icount, node_total, vflags);
      LB_STEP(hipGetLastError());
      hipLaunchKernelGGL(k_gather, dim3(blocks_for(n, 256)), dim3(256), 0, stream, d_pos, d_col, d_mode, vin, n,
                         out->tris, out->cols);
      LB_STEP(hipGetLastError());
      uint32_t cb[6];
      uint32_t bad = 0;
      LB_STEP(hipMemcpyAsync(&host_sb, stack_bound, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
      LB_STEP(hipMemcpyAsync(cb, cbounds, sizeof(cb), hipMemcpyDeviceToHost, stream));
      LB_STEP(hipMemcpyAsync(&bad, vflags, sizeof(bad), hipMemcpyDeviceToHost, stream));
      LB_STEP(hipStreamSynchronize(stream));
      if (bad) {
        if (err) *err = structure_error(bad);
        break;
      }
      for (int a = 0; a < 3; ++a) {
        out->centroid_lo[a] = fkey_inv(cb[a]);
        out->centroid_hi[a] = fkey_inv(cb[3 + a]);
      }
      fail = false;
    } while (0);
    if (fail) break;
    ok = true;
#undef LB_STEP
  } while (0);
  for (void* p : {(void*)tb, (void*)cen, (void*)cbounds, (void*)codes_a, (void*)codes_b, (void*)ids_a,
                  (void*)ids_b, (void*)hist, (void*)left, (void*)right, (void*)parent, (void*)range,
                  (void*)bounds, (void*)cpart, (void*)dist, (void*)vflags, (void*)srcs_a, (void*)srcs_b, (void*)depth, (void*)stack_bound,
                  (void*)icount, (void*)ioffs})
    if (p) (void)hipFree(p);
  scan_release(&scan);
  if (!ok) {
    free_device_scene(out);
    return false;
  }
  out->node_count = node_total;
  out->stack_bound = host_sb;
  out->levels = level_count;
  out->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return true;
}

}  // namespace eye
