#include "batch.h"

#include <algorithm>
#include <cstring>

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

bool same_launch(const UnitLaunch& a, const UnitLaunch& b) {
  return a.omms == b.omms && a.omm_count == b.omm_count && a.poses == b.poses && a.pose_count == b.pose_count &&
         a.pose_base == b.pose_base && a.seed == b.seed && a.eye == b.eye && a.frame == b.frame && a.spp == b.spp &&
         a.omm_counts == b.omm_counts && a.unit_counts == b.unit_counts && a.sample_base == b.sample_base &&
         a.order == b.order && a.colours == b.colours && a.stats == b.stats && a.samples == b.samples &&
         a.sample_offsets == b.sample_offsets && a.queue == b.queue && a.units_per_grab == b.units_per_grab;
}
}  // namespace

void BatchRenderer::release() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
  if (d_poses_) (void)hipFree(d_poses_);
  if (d_colours_) (void)hipFree(d_colours_);
  if (d_queue_) (void)hipFree(d_queue_);
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  d_poses_ = nullptr;
  d_colours_ = nullptr;
  d_queue_ = nullptr;
  ev0_ = ev1_ = nullptr;
  poses_cap_ = colours_cap_ = 0;
  g_blocks_ = -1;
  g_scene_ = nullptr;
}

bool BatchRenderer::build_graph(const DeviceScene& scene, const UnitLaunch& U, int blocks, hipStream_t stream,
                                std::string* err) {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
  hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    if (err) *err = std::string("batch: begin capture: ") + hipGetErrorString(e);
    return false;
  }
  std::string lerr;
  const bool ok = launch_render_units(scene, U, blocks, stream, &lerr);
  e = hipStreamEndCapture(stream, &graph_);
  if (!ok || e != hipSuccess) {
    if (err) *err = !ok ? lerr : std::string("batch: end capture: ") + hipGetErrorString(e);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_ = nullptr;
    return false;
  }
  e = hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    if (err) *err = std::string("batch: instantiate: ") + hipGetErrorString(e);
    return false;
  }
  g_launch_ = U;
  g_blocks_ = blocks;
  g_scene_ = scene.nodes;
  return true;
}

bool BatchRenderer::render(const DeviceScene& scene, UnitLaunch U, const std::vector<Pose>& poses, float* rgb,
                           hipStream_t stream, double* gpu_ms, std::string* err) {
  const size_t np = poses.size();
  *gpu_ms = 0.0;
  if (np == 0) return true;
  if (!ev0_ && (hipEventCreate(&ev0_) != hipSuccess || hipEventCreate(&ev1_) != hipSuccess)) {
    if (err) *err = "batch: event creation failed";
    return false;
  }
  const size_t units = np * (size_t)U.omm_count;
  const bool realloc = poses_cap_ < np || colours_cap_ < units;
  if (!grow(&d_poses_, &poses_cap_, np, err)) return false;
  if (!grow(&d_colours_, &colours_cap_, units, err)) return false;
  if (!d_queue_ && hipMalloc(&d_queue_, sizeof(uint32_t) * 4) != hipSuccess) {
    if (err) *err = "batch: queue allocation failed";
    return false;
  }
  if (hipMemcpyAsync(d_poses_, poses.data(), sizeof(Pose) * np, hipMemcpyHostToDevice, stream) != hipSuccess) {
    if (err) *err = "batch: pose upload failed";
    return false;
  }
  U.poses = d_poses_;
  U.pose_count = (int)np;
  U.pose_base = 0;
  U.colours = d_colours_;
  U.queue = d_queue_;
  const int blocks = 0;
  if (use_graph_) {
    if (realloc || !exec_ || g_scene_ != scene.nodes || !same_launch(U, g_launch_) || g_blocks_ != blocks)
      if (!build_graph(scene, U, blocks, stream, err)) return false;
    (void)hipEventRecord(ev0_, stream);
    if (hipGraphLaunch(exec_, stream) != hipSuccess) {
      if (err) *err = "batch: graph launch failed";
      return false;
    }
  } else {
    (void)hipEventRecord(ev0_, stream);
    if (!launch_render_units(scene, U, blocks, stream, err)) return false;
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
    std::vector<float4> tmp(units);
    if (hipMemcpy(tmp.data(), d_colours_, sizeof(float4) * units, hipMemcpyDeviceToHost) != hipSuccess) {
      if (err) *err = "batch: readback failed";
      return false;
    }
    for (size_t i = 0; i < units; ++i) {
      rgb[3 * i] = tmp[i].x;
      rgb[3 * i + 1] = tmp[i].y;
      rgb[3 * i + 2] = tmp[i].z;
    }
  }
  return true;
}

}  // namespace eye
