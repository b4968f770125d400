// k nearest ommatidia by a 6-D k-d tree (median splits on the axis of largest
// spread, buckets of 8), queried on all host threads.  Deterministic: ties in
// distance are broken by index, and each query's result does not depend on the
// tree's shape.
#include "knn.h"

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace eye {
namespace {

constexpr int kDim = 6;
constexpr int kBucket = 8;

struct KdNode {
  int lo, hi;      // range of the permutation
  int axis;        // -1: bucket
  double split;
  int left, right;
};

struct KdTree {
  const double* pts = nullptr;  // n x kDim
  std::vector<uint32_t> perm;
  std::vector<KdNode> nodes;

  int build(int lo, int hi) {
    const int id = (int)nodes.size();
    nodes.push_back(KdNode{lo, hi, -1, 0.0, -1, -1});
    if (hi - lo <= kBucket) return id;
    double mn[kDim], mx[kDim];
    for (int d = 0; d < kDim; ++d) {
      mn[d] = INFINITY;
      mx[d] = -INFINITY;
    }
    for (int t = lo; t < hi; ++t)
      for (int d = 0; d < kDim; ++d) {
        const double v = pts[(size_t)perm[t] * kDim + d];
        mn[d] = std::min(mn[d], v);
        mx[d] = std::max(mx[d], v);
      }
    int ax = 0;
    for (int d = 1; d < kDim; ++d)
      if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
    if (!(mx[ax] - mn[ax] > 0.0)) return id;  // coincident points: one bucket
    const int mid = lo + (hi - lo) / 2;
    std::nth_element(perm.begin() + lo, perm.begin() + mid, perm.begin() + hi, [&](uint32_t a, uint32_t b) {
      const double va = pts[(size_t)a * kDim + ax], vb = pts[(size_t)b * kDim + ax];
      return va < vb || (va == vb && a < b);
    });
    const double split = pts[(size_t)perm[mid] * kDim + ax];
    const int l = build(lo, mid);
    const int r = build(mid, hi);
    nodes[id].axis = ax;
    nodes[id].split = split;
    nodes[id].left = l;
    nodes[id].right = r;
    return id;
  }
};

struct Best {
  int k = 0, count = 0;
  double d2[16];
  uint32_t j[16];
  bool better(double d, uint32_t idx, int slot) const { return d < d2[slot] || (d == d2[slot] && idx < j[slot]); }
  void offer(double d, uint32_t idx) {
    if (count == k && !better(d, idx, k - 1)) return;
    int s = count < k ? count++ : k - 1;
    while (s > 0 && better(d, idx, s - 1)) {
      d2[s] = d2[s - 1];
      j[s] = j[s - 1];
      --s;
    }
    d2[s] = d;
    j[s] = idx;
  }
  double worst() const { return count < k ? INFINITY : d2[k - 1]; }
};

void query(const KdTree& T, int node, const double* q, uint32_t self, Best& best) {
  const KdNode& N = T.nodes[(size_t)node];
  if (N.axis < 0) {
    for (int t = N.lo; t < N.hi; ++t) {
      const uint32_t j = T.perm[(size_t)t];
      if (j == self) continue;
      const double* p = T.pts + (size_t)j * kDim;
      double d = 0.0;
      for (int a = 0; a < kDim; ++a) {
        const double e = q[a] - p[a];
        d = d + e * e;
      }
      best.offer(d, j);
    }
    return;
  }
  const double diff = q[N.axis] - N.split;
  const int near = diff < 0.0 ? N.left : N.right;
  const int far = diff < 0.0 ? N.right : N.left;
  query(T, near, q, self, best);
  // points equal to the split lie on both sides; <= keeps index ties exact
  if (diff * diff <= best.worst()) query(T, far, q, self, best);
}

}  // namespace

void nearest_ommatidia(const float* omm, size_t n, int k, double pos_scale, uint32_t* out) {
  if (n == 0 || k <= 0) return;
  k = std::min(k, 16);
  std::vector<double> pts(n * kDim);
  for (size_t i = 0; i < n; ++i) {
    const float* o = omm + 8 * i;
    const double dx = o[3], dy = o[4], dz = o[5];
    const double len = std::sqrt(dx * dx + dy * dy + dz * dz);
    double* p = pts.data() + i * kDim;
    if (len > 0.0 && std::isfinite(len)) {
      p[0] = dx / len;
      p[1] = dy / len;
      p[2] = dz / len;
    } else {
      p[0] = p[1] = p[2] = 0.0;
    }
    p[3] = (double)o[0] / pos_scale;
    p[4] = (double)o[1] / pos_scale;
    p[5] = (double)o[2] / pos_scale;
  }
  KdTree T;
  T.pts = pts.data();
  T.perm.resize(n);
  for (size_t i = 0; i < n; ++i) T.perm[i] = (uint32_t)i;
  T.nodes.reserve(2 * (n / kBucket + 1));
  T.build(0, (int)n);
  const int kk = (int)std::min<size_t>((size_t)k, n - 1);
  auto work = [&](size_t a, size_t b) {
    for (size_t i = a; i < b; ++i) {
      Best best;
      best.k = kk;
      if (kk > 0) query(T, 0, pts.data() + i * kDim, (uint32_t)i, best);
      for (int m = 0; m < k; ++m) out[i * (size_t)k + (size_t)m] = m < best.count ? best.j[m] : (uint32_t)i;
    }
  };
  const size_t threads = std::max<size_t>(1, std::min<size_t>(16, std::thread::hardware_concurrency()));
  if (n < 4096 || threads == 1) {
    work(0, n);
    return;
  }
  std::vector<std::thread> pool;
  const size_t per = (n + threads - 1) / threads;
  for (size_t t = 0; t < threads; ++t) {
    const size_t a = t * per, b = std::min(n, a + per);
    if (a < b) pool.emplace_back(work, a, b);
  }
  for (std::thread& th : pool) th.join();
}

}  // namespace eye
