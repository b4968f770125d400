// Render kernels of the compound-eye path (host-side launch interface).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_scene.h"
#include "eyecore.h"

namespace eye {

// One eye rendered at one or more poses.  Sample s of ommatidium i at pose p
// is ray number  p * rays_per_pose + offsets[i] + s  (offsets == nullptr means
// a uniform `spp` per ommatidium, offsets[i] = i * spp).
struct EyeLaunch {
  const Ommatidium* omms = nullptr;
  int omm_count = 0;
  const uint32_t* offsets = nullptr;  // omm_count + 1 entries, or nullptr
  int spp = 1;
  uint64_t rays_per_pose = 0;
  const Pose* poses = nullptr;        // pose_count entries (device)
  int pose_count = 1;
  uint32_t pose_base = 0;             // counter word 3 of the first pose
  uint32_t seed = 42;
  uint32_t eye = 0;
  uint32_t frame = 0;
};

struct TraceStats {
  double ms = 0.0;
};

// Per-sample radiance: samples[ray] = (r, g, b, t) with t < 0 for a miss.
bool launch_trace_samples(const DeviceScene& scene, const EyeLaunch& L, float4* samples,
                          hipStream_t stream, std::string* err);

// Traversal work of the rays of L: totals[0] += nodes visited, totals[1] +=
// triangles tested, totals[2] += rays that hit (device counters, zeroed by the caller).
bool launch_trace_stats(const DeviceScene& scene, const EyeLaunch& L, unsigned long long* totals,
                        hipStream_t stream, std::string* err);

// Per-ommatidium averages in the canonical order (lane l of a wave sums
// samples l, l+64, ... sequentially, then an xor butterfly over the wave):
// colours[p * omm_count + i] = (r, g, b, n).  `stats`, when not null, gets
// (depth mean, depth variance, rgb variance mean, hit fraction) per
// ommatidium over the first `stat_samples` samples.
bool launch_reduce(const EyeLaunch& L, const float4* samples, float4* colours, hipStream_t stream,
                   std::string* err);

// uchar4 frame of the single_dimension_fast projection (shaders.cu:396-406):
// pixel x of row 0 gets make_color(colour[x]) for x < min(width, omm_count).
bool launch_frame_single_dim_fast(const float4* colours, int omm_count, uchar4* frame, int width,
                                  hipStream_t stream, std::string* err);
// single_dimension (shaders.cu:375-390): every pixel (x, y) shows ommatidium
// x * omm_count / width.
bool launch_frame_single_dim(const float4* colours, int omm_count, uchar4* frame, int width, int height,
                             hipStream_t stream, std::string* err);
// raw_ommatidial_samples (shaders.cu:354-369): pixel (x = omm, y = sample).
bool launch_frame_raw_samples(const float4* samples, int omm_count, int spp, uchar4* frame, int width,
                              int height, hipStream_t stream, std::string* err);

// Rays of one eye at one pose, in (ommatidium, sample) order, 8 floats each.
bool launch_generate_rays(const EyeLaunch& L, Ray* rays, hipStream_t stream, std::string* err);

// Closest hits of arbitrary rays: hit[i] = original triangle id or -1,
// tbb[i] = (t, b1, b2, 0).
bool launch_trace_rays(const DeviceScene& scene, const Ray* rays, uint64_t n, int32_t* hit,
                       float4* tbb, hipStream_t stream, std::string* err);

}  // namespace eye
