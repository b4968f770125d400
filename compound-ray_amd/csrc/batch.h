// Batched rendering of one eye at many poses (no reference counterpart: the
// reference renders one pose per renderFrame call, libEyeRenderer.cpp:229-242).
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
  // Renders L (its omms/offsets/spp/seed/eye/frame) at every pose of `poses`;
  // rgb receives 3 floats per (pose, ommatidium), pose-major (may be null).
  // *gpu_ms is the device time of the render kernels.
  bool render(const DeviceScene& scene, EyeLaunch L, const std::vector<Pose>& poses, float* rgb,
              hipStream_t stream, double* gpu_ms, std::string* err);
  void release();

 private:
  Pose* d_poses_ = nullptr;
  size_t poses_cap_ = 0;
  float4* d_samples_ = nullptr;
  size_t samples_cap_ = 0;
  float4* d_colours_ = nullptr;
  size_t colours_cap_ = 0;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
};

}  // namespace eye
