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

// Launch shape of the unit renderer.
struct LaunchShape {
  int units_per_grab = 4;
  int blocks = 0;  // 0: resident capacity
  int traversal = 0;  // UnitTraversal
};

struct TuneTrial {
  LaunchShape shape;
  double ms = 0.0;  // fastest of the timed launches
};

// Launch-shape autotuner: candidate shapes are timed on a sample of a batch and
// the fastest is kept per workload key (scene, eye and sample-count totals).
// Entries can be saved to and loaded from a text file, one per line:
//   <key> <units_per_grab> <blocks> <traversal> <ms>
class LaunchTuner {
 public:
  // grabs {1,2,4,8,16} x blocks {capacity/2, capacity, 2 capacity} x traversal {packet, packet with
  // scalar loads, per-ray}; capacity is that of the packet kernel
  static std::vector<LaunchShape> candidates(int capacity);
  static std::string key_of(const DeviceScene& scene, const UnitLaunch& U, uint64_t samples_per_pose);
  bool find(const std::string& key, LaunchShape* shape) const;
  void put(const std::string& key, const LaunchShape& shape, double ms);
  bool load(const std::string& path, std::string* err);
  bool save(const std::string& path, std::string* err) const;
  size_t size() const { return keys_.size(); }

 private:
  std::vector<std::string> keys_;
  std::vector<LaunchShape> shapes_;
  std::vector<double> ms_;
};

class BatchRenderer {
 public:
  ~BatchRenderer() { release(); }
  // Renders U (omms, counts, order, seed, eye, frame, launch shape) at every
  // pose of `poses`; rgb receives 3 floats per (pose, ommatidium), pose-major
  // (may be null).  *gpu_ms is the device time of the render launch.
  bool render(const DeviceScene& scene, UnitLaunch U, const std::vector<Pose>& poses, float* rgb,
              hipStream_t stream, double* gpu_ms, std::string* err);
  // Times every candidate shape on `poses` (each launched once untimed, then
  // `repeats` times); trials receives one entry per candidate in order, and
  // *best the fastest.  Colours of the last trial stay in colours().
  bool tune(const DeviceScene& scene, UnitLaunch U, const std::vector<Pose>& poses, int repeats,
            hipStream_t stream, std::vector<TuneTrial>* trials, LaunchShape* best, std::string* err);
  // Device colours of the last batch, [pose][omm] float4.
  const float4* colours() const { return d_colours_; }
  void set_use_graph(bool on) { use_graph_ = on; }
  void release();

 private:
  bool build_graph(const DeviceScene& scene, const UnitLaunch& U, hipStream_t stream, std::string* err);
  bool prepare(const UnitLaunch& U, const std::vector<Pose>& poses, hipStream_t stream, bool* realloc,
               std::string* err);
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
  uint64_t g_scene_ = 0;  // generation of the scene the graph was captured on
  UnitLaunch g_launch_;
  bool g_valid_ = false;
};

}  // namespace eye
