// Learned adaptive sampling (no reference counterpart; the reference samples
// every ommatidium uniformly and picks one rate per scene,
// python-examples/minimum-samples-calculation/readme.txt).
//
// Per (pose, ommatidium) unit:
//   1. pilot: `pilot` samples with statistics (mean, per-channel variance,
//      hit fraction, inverse-depth mean and variance);
//   2. features (16): world-space axis (3), acceptance angle / 0.05,
//      pilot mean RGB (3), pilot RGB standard deviation (3), mean inverse
//      depth 1/(1+t) (0 for a miss), its standard deviation; and over the
//      unit's kNeighbours nearest ommatidia (same pose): the mean and the
//      maximum of their pilot standard deviation (RGB mean), the mean
//      absolute difference of their pilot mean RGB from the unit's, and the
//      mean of their inverse-depth standard deviation;
//   3. the MLP predicts log2 of the samples the unit needs for an RMS error of
//      `error_target` on its own: n_req = mean RGB variance / error^2 (or,
//      without the network, n_req from the pooled pilot deviation, below);
//   4. allocation minimising the total samples for a mean squared error of
//      error_target^2 over the eye (Neyman): n_i = sqrt(n_req,i) * sum_j
//      sqrt(n_req,j) / N, clamped to [pilot, max_spp];
//   5. top-up: samples pilot .. n_i-1 of every unit (the same random streams a
//      uniform render would use), combined with the pilot by sample count.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "device_scene.h"
#include "knn.h"
#include "mlp.h"
#include "render.h"

namespace eye {

struct AdaptiveConfig {
  float error_target = 0.01f;
  float label_error = 0.01f;   // error the network's labels were made for: its
                               //   n_req scale by (label_error / error_target)^2
  int pilot = 8;
  int max_spp = 512;
  float budget_scale = 1.0f;   // multiplies every allocation (calibration knob)
  const uint32_t* neighbours = nullptr;  // device, kNeighbours per ommatidium (nearest_ommatidia)
  int sampler = 0;             // AdaptiveSampler: where n_req comes from
};

// Source of each unit's required samples n_req: the trained network, or the
// network-free estimate from the pooled pilot standard deviation of the unit
// and its kNeighbours nearest ommatidia, sigma = (own + k * neighbours' mean)
// / (k + 1), n_req = sigma^2 / label_error^2.  The allocation, the top-up and
// the combination are the same for both.
enum AdaptiveSampler : int { kSamplerNetwork = 0, kSamplerPooledSigma = 1 };

struct AdaptiveResult {
  double gpu_ms = 0.0;
  uint64_t total_samples = 0;
  // device milliseconds of the stages: pilot render, features + n_req +
  // allocation, top-up render, combination
  double stage_ms[4] = {0.0, 0.0, 0.0, 0.0};
};

class AdaptiveRenderer {
 public:
  ~AdaptiveRenderer() { release(); }
  // Renders U's eye at every pose with learned allocation; rgb (nullable)
  // gets 3 floats per (pose, ommatidium), counts (nullable) the samples used.
  bool render(const DeviceScene& scene, UnitLaunch U, const std::vector<Pose>& poses, Mlp& mlp,
              const AdaptiveConfig& cfg, float* rgb, uint32_t* counts, AdaptiveResult* res, hipStream_t stream,
              std::string* err);
  // Training rows: features of `pilot` samples and labels log2(n_req) from
  // `label_spp` further samples of the chosen units (unit_mask[pose][omm] != 0).
  bool training_set(const DeviceScene& scene, UnitLaunch U, const std::vector<Pose>& poses,
                    const std::vector<uint8_t>& unit_mask, int pilot, int label_spp, float error_target,
                    const uint32_t* neighbours, std::vector<float>* features, std::vector<float>* labels, hipStream_t stream, std::string* err);
  // Pilot and top-up colours (r, g, b, samples) and top-up counts of the
  // first `units` units of the last render.
  bool copy_last(size_t units, float4* pilot, float4* topup, uint32_t* counts, std::string* err);
  void release();

 private:
  bool ensure(size_t poses, size_t units, std::string* err);
  Pose* d_poses_ = nullptr;
  size_t poses_cap_ = 0;
  float4* d_col_a_ = nullptr;     // pilot colours
  float4* d_col_b_ = nullptr;     // top-up colours
  float4* d_final_ = nullptr;     // combined colours
  float4* d_stats_ = nullptr;     // pilot statistics (2 per unit)
  float* d_feat_ = nullptr;       // kMlpIn per unit
  float* d_pred_ = nullptr;       // per unit
  uint32_t* d_counts_ = nullptr;  // final sample count per unit
  uint32_t* d_extra_ = nullptr;   // top-up samples per unit
  unsigned long long* d_total_ = nullptr;
  float* d_part_ = nullptr;       // partial sums of the allocation
  size_t part_cap_ = 0;
  float* d_sqrt_ = nullptr;       // per pose: sum of sqrt(n_req)
  uint32_t* d_queue_ = nullptr;
  size_t units_cap_ = 0;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  hipEvent_t ev_stage_[3] = {nullptr, nullptr, nullptr};   // after the pilot, the allocation, the top-up
};

}  // namespace eye
