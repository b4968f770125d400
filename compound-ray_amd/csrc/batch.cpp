#include "batch.h"

#include <algorithm>

namespace eye {

namespace {
template <typename T>
bool grow(T** p, size_t* cap, size_t n, std::string* err) {
  if (*cap >= n && *p) return true;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  if (n == 0) return true;
  const hipError_t e = hipMalloc(p, sizeof(T) * n);
  if (e != hipSuccess) {
    if (err) *err = std::string("batch: hipMalloc failed: ") + hipGetErrorString(e);
    return false;
  }
  *cap = n;
  return true;
}
}  // namespace

void BatchRenderer::release() {
  if (d_poses_) (void)hipFree(d_poses_);
  if (d_samples_) (void)hipFree(d_samples_);
  if (d_colours_) (void)hipFree(d_colours_);
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  d_poses_ = nullptr;
  d_samples_ = nullptr;
  d_colours_ = nullptr;
  ev0_ = ev1_ = nullptr;
  poses_cap_ = samples_cap_ = colours_cap_ = 0;
}

bool BatchRenderer::render(const DeviceScene& scene, EyeLaunch L, const std::vector<Pose>& poses, float* rgb,
                           hipStream_t stream, double* gpu_ms, std::string* err) {
  const size_t np = poses.size();
  if (np == 0) {
    *gpu_ms = 0.0;
    return true;
  }
  if (!ev0_ && (hipEventCreate(&ev0_) != hipSuccess || hipEventCreate(&ev1_) != hipSuccess)) {
    if (err) *err = "batch: event creation failed";
    return false;
  }
  // poses per chunk so that the per-sample buffer stays below 2^28 entries
  const uint64_t rpp = std::max<uint64_t>(L.rays_per_pose, 1);
  const size_t chunk = (size_t)std::max<uint64_t>(1, std::min<uint64_t>(np, (1ull << 28) / rpp));
  if (!grow(&d_poses_, &poses_cap_, np, err)) return false;
  if (!grow(&d_samples_, &samples_cap_, (size_t)(rpp * chunk), err)) return false;
  if (!grow(&d_colours_, &colours_cap_, (size_t)L.omm_count * np, err)) return false;
  if (hipMemcpyAsync(d_poses_, poses.data(), sizeof(Pose) * np, hipMemcpyHostToDevice, stream) != hipSuccess) {
    if (err) *err = "batch: pose upload failed";
    return false;
  }
  (void)hipEventRecord(ev0_, stream);
  for (size_t p0 = 0; p0 < np; p0 += chunk) {
    EyeLaunch C = L;
    C.poses = d_poses_ + p0;
    C.pose_count = (int)std::min(chunk, np - p0);
    C.pose_base = (uint32_t)p0;
    if (!launch_trace_samples(scene, C, d_samples_, stream, err)) return false;
    if (!launch_reduce(C, d_samples_, d_colours_ + p0 * (size_t)L.omm_count, stream, err)) return false;
  }
  (void)hipEventRecord(ev1_, stream);
  if (hipEventSynchronize(ev1_) != hipSuccess) {
    if (err) *err = "batch: render failed";
    return false;
  }
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, ev0_, ev1_);
  *gpu_ms = ms;
  if (rgb) {
    std::vector<float4> tmp((size_t)L.omm_count * np);
    if (hipMemcpy(tmp.data(), d_colours_, sizeof(float4) * tmp.size(), hipMemcpyDeviceToHost) != hipSuccess) {
      if (err) *err = "batch: readback failed";
      return false;
    }
    for (size_t i = 0; i < tmp.size(); ++i) {
      rgb[3 * i] = tmp[i].x;
      rgb[3 * i + 1] = tmp[i].y;
      rgb[3 * i + 2] = tmp[i].z;
    }
  }
  return true;
}

}  // namespace eye
