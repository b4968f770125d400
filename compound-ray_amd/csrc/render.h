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

// Long-task split of the packet task kernel (see UnitLaunch::split_cost).  The
// tasks listed from the last launch's packet steps are traced before every
// other task, in three tiers by length: the longest as 4 packets of 16 rays,
// the next as 2 packets of 32 rays, the next whole (started early).
constexpr int kSplit4Max = 128;             // tier 0: tasks traced as 4 packets, at most
constexpr int kSplit2Max = 896;             // tier 1: tasks traced as 2 packets, at most
constexpr int kSplit1Max = 2048;            // tier 2: tasks started early, whole, at most
constexpr uint32_t kSplitMinSteps = 128;    // tiers 0 and 1: a task of fewer packet steps is never split
constexpr uint32_t kSplitEarlyMinSteps = 64;  // tier 2: nor started early
constexpr int kSplitListMax = kSplit4Max + kSplit2Max + kSplit1Max;
constexpr int kSplitSubMax = kSplit4Max + kSplit2Max;   // listed tasks traced as several packets
constexpr int kSplitSubWaves = 4 * kSplit4Max + 2 * kSplit2Max + kSplit1Max;
constexpr int kSplitCountWords = 8;         // listed per tier, then the tiers' thresholds in steps
constexpr int kSplitBins = 256;             // threshold search: histogram bins of 8 steps
// Device words of a split buffer for `tasks` tasks: cost and flag per task,
// the tiers' lists, the counts and thresholds, a completion counter per
// split task.
constexpr uint64_t split_buffer_words(uint64_t tasks) {
  return 2 * tasks + (uint64_t)kSplitListMax + kSplitCountWords + kSplitSubMax;
}

// Persistent ommatidium-unit renderer.  A unit is one ommatidium at one pose;
// waves take units from a work queue (`units_per_grab` at a time), trace a
// unit's samples 64 at a time as one wave packet and reduce them in registers
// in the canonical order (lane l sums samples l, l+64, ... then an xor
// butterfly), so no per-sample buffer is needed.
struct UnitLaunch {
  const Ommatidium* omms = nullptr;
  int omm_count = 0;
  const Pose* poses = nullptr;
  int pose_count = 1;
  uint32_t pose_base = 0;
  uint32_t seed = 42, eye = 0, frame = 0;
  int spp = 1;                          // uniform count when both count arrays are null
  const uint32_t* omm_counts = nullptr;   // per ommatidium, shared by every pose
  const uint32_t* unit_counts = nullptr;  // per (pose, ommatidium), pose-major
  uint32_t sample_base = 0;             // sample index of each unit's first sample
  const uint32_t* order = nullptr;      // processing order of the ommatidia (null: index order)
  float4* colours = nullptr;            // [pose][omm] = (r, g, b, n)
  float4* stats = nullptr;              // nullable, [pose][omm][2]: (var r, var g, var b, hit fraction),
                                        //   (mean inverse depth, its variance, 0, 0)
  float4* samples = nullptr;            // nullable: per-sample (r, g, b, t or -1)
  const uint64_t* sample_offsets = nullptr;  // first sample record of each unit
  uint32_t* queue = nullptr;            // device work counters (kQueueWords)
  // Optional scratch for split launches (uniform sample counts and few
  // units): each unit's 64-sample chunks go to different waves; per-lane
  // chunk values (1 float4, 3 with stats) and a completion counter per unit.
  float4* chunk_partial = nullptr;
  uint64_t chunk_partial_cap = 0;       // float4 elements
  uint32_t* chunk_done = nullptr;
  uint64_t chunk_done_cap = 0;          // counters
  // Optional long-task split of the packet task kernel (an eye rendered frame
  // after frame, one pose, uniform counts).  The packet steps each task took
  // in the last launch (split_cost, per task) predict this launch's longest
  // tasks, which are listed (split_list, three tiers of kSplit4Max,
  // kSplit2Max and kSplit1Max entries) and flagged (split_flag: tier + 1) and
  // traced first, split into packets of rays of neighbouring elevation or
  // whole.  Results do not depend on it.  split_count: kSplitCountWords;
  // split_sub_done: kSplitSubMax.  split_plan: split_cost holds the steps of
  // a launch of the same tasks.
  uint32_t* split_cost = nullptr;
  uint32_t* split_flag = nullptr;
  uint32_t* split_list = nullptr;
  uint32_t* split_count = nullptr;      // [0..2] tasks listed per tier, [3..5] the tiers' thresholds
  uint32_t* split_sub_done = nullptr;
  uint64_t split_tasks = 0;
  bool split_plan = false;
  // launch shape (results do not depend on it)
  int units_per_grab = 0;               // consecutive units a wave takes from the queue (0: chosen
                                        //   from the launch size, at most auto_grab_max)
  int auto_grab_max = 16;
  int blocks = 0;                       // blocks of 256 threads; 0: every resident slot once
  int traversal = 0;                    // UnitTraversal
  // nullable: packet traversals add their node visits ([0]) and the triangles
  // of the leaves they enter ([1]), counted once per wave (measurement runs)
  unsigned long long* packet_counts = nullptr;
};

// Traversal inside the unit renderer: the wave walks the BVH as one packet
// (a wave-uniform node, node records read with vector loads or through the
// scalar unit), or every lane walks its own ray with an LDS stack per lane.
enum UnitTraversal : int { kUnitPacket = 0, kUnitPacketScalar = 1, kUnitPerRay = 2 };

// Work counters a launch needs at U.queue: one per slice of the unit range.
constexpr int kQueueWords = 8;
// Largest grab of the unit renderer: lane g of a wave holds unit g's count.
constexpr int kMaxUnitsPerGrab = 64;

// Launches the unit renderer with U's launch shape.  The queue counter is
// cleared by a kernel first.
bool launch_render_units(const DeviceScene& scene, const UnitLaunch& U, hipStream_t stream, std::string* err);

// Name of the kernel the last launch_render_units used (k_render_units or the
// split-launch k_render_chunks), and the switch for split launches (on by default).
const char* last_render_kernel();
void set_split_launches(bool on);

// Blocks of the unit renderer that fit on the device at once (CUs x resident
// blocks per CU for this scene's stack size and U's kernel variant).
int render_units_capacity(const DeviceScene& scene, const UnitLaunch& U);

// Traversal of eye rays in renderFrame and the batch paths: 0 = one ray per
// lane, 1 = wave packets, 2 = wave packets reading nodes and triangles
// through the scalar unit (UnitTraversal kUnitPerRay / kUnitPacket /
// kUnitPacketScalar).
enum TraceMode : int { kTracePerRay = 0, kTracePacket = 1, kTracePacketScalar = 2 };
void set_trace_mode(int mode);
int trace_mode();
int unit_traversal_of(int trace_mode);
int trace_mode_of(int unit_traversal);

// Plain float4 copy of n elements (bandwidth reference for the profiler).
bool launch_copy_probe(const float4* src, float4* dst, size_t n, hipStream_t stream, std::string* err);

// Traversal work of the rays of L: totals[0] += nodes visited, totals[1] +=
// triangles tested, totals[2] += rays that hit (device counters, zeroed by the caller).
bool launch_trace_stats(const DeviceScene& scene, const EyeLaunch& L, unsigned long long* totals,
                        hipStream_t stream, std::string* err);

// uchar4 frame of the single_dimension_fast projection (shaders.cu:396-406):
// pixel x of row 0 gets make_color(colour[x]) for x < min(width, omm_count).
bool launch_frame_single_dim_fast(const float4* colours, int omm_count, uchar4* frame, int width,
                                  hipStream_t stream, std::string* err);
// single_dimension (shaders.cu:375-390): every pixel (x, y) shows ommatidium
// x * omm_count / width.
bool launch_frame_single_dim(const float4* colours, int omm_count, uchar4* frame, int width, int height,
                             hipStream_t stream, std::string* err);
// raw_ommatidial_samples (shaders.cu:354-369): pixel (x = omm, y = sample).
// offsets: prefix sum of per-ommatidium sample counts (omm_count + 1 entries),
// or null for spp samples each; ommatidium x fills rows 0 .. count_x - 1.
bool launch_frame_raw_samples(const float4* samples, const uint32_t* offsets, int omm_count, int spp,
                              uchar4* frame, int width, int height, hipStream_t stream, std::string* err);

// Frame of a regular camera (pinhole, orthographic or panoramic,
// shaders.cu:198-333): every pixel traces its primary ray and shows
// make_color of the hit's shading or the sky.  frame and rays are nullable.
bool launch_camera_frame(const DeviceScene& scene, const CameraLaunch& C, uchar4* frame, Ray* rays,
                         hipStream_t stream, std::string* err);

// Rays of one eye at one pose, in (ommatidium, sample) order, 8 floats each.
bool launch_generate_rays(const EyeLaunch& L, Ray* rays, hipStream_t stream, std::string* err);

// Closest hits of arbitrary rays: hit[i] = original triangle id or -1,
// tbb[i] = (t, b1, b2, 0).
// One wave traces rays[0..63] with the packet traversal; lanes at or above
// active_lanes skip the call, so below 64 the traversal's partial-wave guard
// sets the scene's fault word.  ids[lane]: hit id, -1 miss, -2 skipped.
bool launch_packet_guard_probe(const DeviceScene& scene, const Ray* rays, int active_lanes, int32_t* ids,
                               hipStream_t stream, std::string* err);

bool launch_trace_rays(const DeviceScene& scene, const Ray* rays, uint64_t n, int32_t* hit,
                       float4* tbb, hipStream_t stream, std::string* err);

}  // namespace eye
