// Acceleration-structure format and the deterministic LBVH rules shared by the
// GPU builder (lbvh.hip) and the GPU traversal (render.hip).
//
// Build (replaces MulticamScene::buildMeshAccels / buildInstanceAccel,
// libEyeRenderer3/MulticamScene.cpp:1052-1428, whose OptiX GAS/IAS are opaque):
//   1. every triangle's AABB centroid is quantised to 10 bits per axis against
//      the centroid bounds, in FP64, and bit-interleaved into a 30-bit Morton
//      code;
//   2. (code, triangle id) pairs are sorted (stable LSD radix sort on the code
//      over ids in increasing order);
//   3. Karras' binary radix tree ("Maximizing parallelism in the construction
//      of BVHs, octrees, and k-d trees", HPG 2012) is built over the sorted
//      codes, ties between equal codes broken by index;
//   4. bounds are refitted bottom-up (min/max, order independent, exact);
//   5. the binary tree is collapsed top-down into kBvhWidth-wide nodes:
//      starting from a node's two children, the internal child with the
//      largest surface area (first in slot order on ties) is opened until
//      kBvhWidth children exist or none can be opened; any binary node
//      covering at most `leaf_max` triangles becomes a leaf that references
//      its contiguous range of sorted triangles.  Wide nodes are numbered
//      breadth-first, level by level, in parent order then slot order.
// Every rule is deterministic, so a CPU build of the same codes produces the
// same node array byte for byte.
#pragma once

#include <stdint.h>
#include "detmath.h"

namespace eye {

// Children per wide node: 4 or 8 (a multiple of kHalfWidth).  The build,
// the traversals and the CPU oracle all follow this constant.  Measured on
// MI355X: an 8-wide tree cut the packet traversal's node visits by a third
// (44.7 against 66.1 per 64 rays at bee scale) but, at twice the bytes and
// slab tests per visit, rendered 2-5 % slower on every workload, so the tree
// is 4-wide.
constexpr int kBvhWidth = 4;
constexpr int kHalfWidth = 4;                        // child slots per 128-byte half
constexpr int kBvhHalves = kBvhWidth / kHalfWidth;
static_assert(kBvhWidth % kHalfWidth == 0, "whole node halves");
constexpr int kLeafMaxLimit = 16;
constexpr int32_t kChildEmpty = 0x7fffffff;
// Far-away degenerate box for empty child slots: never intersected by a ray
// with |origin| < 1e30 and tmax <= 1e16.
constexpr float kEmptyBox = 3.0e38f;

// 128 bytes holding four child slots as structure-of-arrays, so that a lane
// (or the scalar unit) reads them as eight aligned 16-byte vectors.
struct alignas(16) WideHalf {
  float lo_x[4], hi_x[4];
  float lo_y[4], hi_y[4];
  float lo_z[4], hi_z[4];
  int32_t child[4];   // >= 0 internal wide node; < 0 leaf; kChildEmpty unused
  uint32_t pad[4];
};
static_assert(sizeof(WideHalf) == 128, "a node half must be 128 bytes");

// Wide node: child slot s lives in half s / 4 at position s % 4.  Slots fill
// in order; empty slots hold kChildEmpty and the far box.
struct alignas(16) WideNode {
  WideHalf half[kBvhHalves];
};
static_assert(sizeof(WideNode) == 128 * kBvhHalves, "wide node layout");

// Leaf child encoding: bit 31 set, bits 27..30 = count-1, bits 0..26 = first
// sorted triangle.
EYE_HD int32_t encode_leaf(uint32_t first, uint32_t count) {
  return (int32_t)(0x80000000u | ((count - 1u) << 27) | first);
}
EYE_HD bool child_is_leaf(int32_t c) { return c < 0; }
EYE_HD uint32_t leaf_first(int32_t c) { return (uint32_t)c & 0x07ffffffu; }
EYE_HD uint32_t leaf_count(int32_t c) { return (((uint32_t)c >> 27) & 0xfu) + 1u; }

// Spreads the low 10 bits of v so that bit i moves to bit 3i.
EYE_HD uint32_t expand_bits10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// FP64 quantisation of one centroid coordinate to [0, 1023].
EYE_HD uint32_t quantise10(float c, float lo, float hi) {
  const double ext = (double)hi - (double)lo;
  if (!(ext > 0.0)) return 0u;
  double q = floor(((double)c - (double)lo) / ext * 1024.0);
  if (q < 0.0) q = 0.0;
  if (q > 1023.0) q = 1023.0;
  return (uint32_t)q;
}

EYE_HD uint32_t morton30(float cx, float cy, float cz, const float lo[3], const float hi[3]) {
  const uint32_t qx = quantise10(cx, lo[0], hi[0]);
  const uint32_t qy = quantise10(cy, lo[1], hi[1]);
  const uint32_t qz = quantise10(cz, lo[2], hi[2]);
  return (expand_bits10(qx) << 2) | (expand_bits10(qy) << 1) | expand_bits10(qz);
}

EYE_HD int clz32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __clz((int)x);
#else
  return x == 0 ? 32 : __builtin_clz(x);
#endif
}

// Common-prefix length of sorted keys i and j (codes tie-broken by index).
EYE_HD int karras_delta(const uint32_t* codes, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = codes[i], b = codes[j];
  if (a != b) return clz32(a ^ b);
  return 32 + clz32((uint32_t)i ^ (uint32_t)j);
}

// Karras 2012, internal node i of n-1.  Children: index, and whether a leaf.
EYE_HD void karras_node(const uint32_t* codes, int n, int i, int* left, bool* left_leaf,
                        int* right, bool* right_leaf, int* first, int* last) {
  const int d = (karras_delta(codes, n, i, i + 1) - karras_delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = karras_delta(codes, n, i, i - d);
  int lmax = 2;
  while (karras_delta(codes, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (karras_delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = karras_delta(codes, n, i, j);
  int s = 0;
  int div = 2;
  for (;;) {
    const int t = (l + div - 1) / div;
    if (karras_delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
    div *= 2;
  }
  const int gamma = i + s * d + (d < 0 ? -1 : 0);
  const int lo = i < j ? i : j;
  const int hi = i < j ? j : i;
  *left = gamma;
  *left_leaf = (lo == gamma);
  *right = gamma + 1;
  *right_leaf = (hi == gamma + 1);
  *first = lo;
  *last = hi;
}

EYE_HD float half_area(const float* b) {  // b = lo x,y,z, hi x,y,z
  const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
  return dx * dy + dy * dz + dz * dx;
}

// Binary tree reference: internal nodes 0..n-2, leaves n-1..2n-2 (leaf k is
// stored at index n-1+k).  `bounds` holds 6 floats per binary node in that
// combined indexing; `range` 2 ints (first, last) per binary node.
struct BinaryTreeView {
  const int32_t* left;    // per internal node, combined index
  const int32_t* right;
  const float* bounds;
  const int32_t* range;
  int n;
};

EYE_HD int bin_size(const BinaryTreeView& t, int node) {
  return t.range[2 * node + 1] - t.range[2 * node] + 1;
}
EYE_HD bool bin_is_leaf(const BinaryTreeView& t, int node) { return node >= t.n - 1; }

// Children of the wide node built from binary node `src`: fills up to eight
// combined binary indices, returns the count.
EYE_HD int collapse_children(const BinaryTreeView& t, int src, int leaf_max, int out[kBvhWidth]) {
  int cnt;
  if (bin_is_leaf(t, src) || bin_size(t, src) <= leaf_max) {
    out[0] = src;
    return 1;
  }
  out[0] = t.left[src];
  out[1] = t.right[src];
  cnt = 2;
  while (cnt < kBvhWidth) {
    int best = -1;
    float best_area = -1.0f;
    for (int k = 0; k < cnt; ++k) {
      const int c = out[k];
      if (bin_is_leaf(t, c) || bin_size(t, c) <= leaf_max) continue;
      const float a = half_area(t.bounds + 6 * c);
      if (a > best_area) {
        best_area = a;
        best = k;
      }
    }
    if (best < 0) break;
    const int c = out[best];
    for (int k = cnt; k > best + 1; --k) out[k] = out[k - 1];
    out[best] = t.left[c];
    out[best + 1] = t.right[c];
    ++cnt;
  }
  return cnt;
}

EYE_HD bool wide_child_is_internal(const BinaryTreeView& t, int c, int leaf_max) {
  return !(bin_is_leaf(t, c) || bin_size(t, c) <= leaf_max);
}

}  // namespace eye
