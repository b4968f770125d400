// Batched rendering of one eye at many poses (no reference counterpart: the
// reference renders one pose per renderFrame call, libEyeRenderer.cpp:229-242).
//
// All (pose, ommatidium) units of a batch go through one persistent launch of
// the unit renderer.  The launch sequence (queue reset + render) is captured
// once into a HIP graph per batch shape and replayed; the launch shape (units
// per queue grab, blocks) comes from the autotuner when one is attached.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "device_scene.h"
#include "render.h"

namespace eye {

class BatchRenderer {
 public:
  ~BatchRenderer() { release(); }
  // Renders U (omms, counts, order, seed, eye, frame) at every pose of
  // `poses`; rgb receives 3 floats per (pose, ommatidium), pose-major (may be
  // null).  *gpu_ms is the device time of the render launch.
  bool render(const DeviceScene& scene, UnitLaunch U, const std::vector<Pose>& poses, float* rgb,
              hipStream_t stream, double* gpu_ms, std::string* err);
  // Device colours of the last batch, [pose][omm] float4.
  const float4* colours() const { return d_colours_; }
  void set_use_graph(bool on) { use_graph_ = on; }
  void release();

 private:
  bool build_graph(const DeviceScene& scene, const UnitLaunch& U, int blocks, hipStream_t stream, std::string* err);
  Pose* d_poses_ = nullptr;
  size_t poses_cap_ = 0;
  float4* d_colours_ = nullptr;
  size_t colours_cap_ = 0;
  uint32_t* d_queue_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  bool use_graph_ = true;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  // shape the captured graph was built for
  const void* g_scene_ = nullptr;
  UnitLaunch g_launch_;
  int g_blocks_ = -1;
};

}  // namespace eye
