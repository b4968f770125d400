// libEyeRenderer3 C ABI on MI355X (include/eye_renderer.h).
//
// Mirrors the global-singleton renderer of libEyeRenderer3/libEyeRenderer.cpp
// (scene, camera list, current camera, render size, frame buffer), minus the
// window: GPU state is created lazily on the first call that needs it, every
// entry point catches its own failures, and nothing throws across the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/eye_renderer.h"
#include "adaptive.h"
#include "batch.h"
#include "device_scene.h"
#include "knn.h"
#include "mlp.h"
#include "noalloc.h"
#include "render.h"
#include "projection.h"
#include "scene_loader.h"

using namespace eye;

namespace {

struct EyeState {
  std::vector<Ommatidium> omms;
  int spp = 1;
  std::vector<uint32_t> counts;  // per-ommatidium sample counts, empty = uniform spp
  std::vector<uint32_t> offsets; // counts' exclusive prefix sum (size omms + 1)
  uint32_t frame = 0;
  uint32_t sample_base = 0;      // random-stream index of each ommatidium's first sample
  uint32_t eye_index = 0;
  std::string data_path;
  std::string shader = "single_dimension_fast";
  Ommatidium* d_omms = nullptr;
  size_t d_omms_cap = 0;
  uint32_t* d_offsets = nullptr;
  size_t d_offsets_cap = 0;
  uint32_t* d_counts = nullptr;          // per-ommatidium sample counts (when counts are set)
  size_t d_counts_cap = 0;
  uint64_t* d_sample_offsets = nullptr;  // first per-sample record of each ommatidium
  size_t d_sample_offsets_cap = 0;
  uint32_t* d_order = nullptr;           // processing order: ommatidia sorted by axis direction
  size_t d_order_cap = 0;
  bool omm_dirty = true;      // ommatidia changed: upload them, recompute the processing order
  bool counts_dirty = true;   // sample counts changed: upload count and sample-record offsets
  bool device_dirty() const { return omm_dirty || counts_dirty; }
  float4* d_samples = nullptr;
  size_t d_samples_cap = 0;
  float4* d_colours = nullptr;
  size_t d_colours_cap = 0;
  bool has_render = false;
  uint64_t rendered_rays = 0;
  // scratch of split renders (per-lane chunk values, per-unit counters)
  float4* d_chunk_partial = nullptr;
  size_t d_chunk_partial_cap = 0;
  uint32_t* d_chunk_done = nullptr;
  size_t d_chunk_done_cap = 0;
  // long-task split of split renders: per task the packet steps of the last
  // render and a flag, the split list and counters (render.h, UnitLaunch)
  uint32_t* d_split = nullptr;
  size_t d_split_cap = 0;
  uint64_t split_tasks = 0;       // tasks the buffers are laid out for
  bool split_valid = false;       // the cost words hold the steps of a render of those tasks
  uint32_t split_listed[3] = {0, 0, 0};   // tasks the last render listed per tier
  uint64_t split_generation = 0;  // scene the costs were measured on
  // automatic choice: renders of the same tasks since the last trial began,
  // and the shortest trace time (ms) of the trial's renders without / with
  uint32_t split_frames = 0;
  float split_ms[2] = {1e30f, 1e30f};
  // pixel -> ommatidium map of the spherical projections, valid for
  // (proj_w, proj_h, proj_variant) and the ommatidia of proj_version
  uint32_t* d_proj_ids = nullptr;
  size_t d_proj_cap = 0;
  int proj_w = -1, proj_h = -1, proj_variant = -1;
  uint64_t proj_version = ~0ull;
  uint64_t omm_version = 0;   // bumped whenever the ommatidia change
  // nearest-ommatidium table of the learned sampler, for the ommatidia of nbr_version
  uint32_t* d_nbr = nullptr;
  size_t d_nbr_cap = 0;
  uint64_t nbr_version = ~0ull;

  uint64_t total_samples() const {
    if (counts.empty()) return (uint64_t)omms.size() * (uint64_t)spp;
    return offsets.empty() ? 0 : offsets.back();
  }
  void restart_streams() { frame = 0; }
};

struct Cam {
  std::string name;
  CameraKind kind = kPerspective;
  Pose pose;
  int eye = -1;
  float yfov_deg = 60.0f;
  // pinhole: (tan(yfov/2), tan(yfov/2), 1), or (10, 10, 1) for the default
  // camera, which never gets a field of view (PerspectiveCamera.cpp:3-7);
  // orthographic: (xmag, ymag, -)
  float scale[3] = {10.0f, 10.0f, 1.0f};
  float start_radius = 0.0f;   // panoramic (PanoramicCamera.cpp:16)
};

struct State {
  bool verbose = true;
  std::string last_error;
  bool hip_ready = false;
  hipStream_t stream = nullptr;
  bool loaded = false;
  HostScene host;
  DeviceScene scene;
  float* d_pos = nullptr;
  float* d_col = nullptr;
  uint32_t* d_mode = nullptr;
  uint32_t* d_texels = nullptr;
  TexView* d_textures = nullptr;
  std::vector<Cam> cams;
  std::vector<EyeState> eyes;
  size_t current = 0;
  int width = 400, height = 400;
  std::vector<uint8_t> frame_host;
  uchar4* d_frame = nullptr;
  size_t d_frame_cap = 0;
  bool frame_host_stale = false;
  Pose* d_pose = nullptr;
  uint32_t seed = 42;
  int leaf_max = 4;
  size_t tri_count = 0;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double timings[4] = {0, 0, 0, 0};   // trace, reduce, projection, total GPU ms of the last render
  unsigned long long* d_stats = nullptr;
  unsigned long long* d_packet_counts = nullptr;  // packet node visits, leaf triangles (when counting)
  bool count_packets = false;
  uint32_t* h_fault = nullptr;   // mapped host word: packet traversal entered by a partial wave
  uint32_t* d_fault = nullptr;   // its device address
  BatchRenderer batch;
  Mlp mlp;
  AdaptiveRenderer adaptive;
  bool record_samples = false;   // keep per-sample records of renderFrame
  uint32_t* d_queue = nullptr;   // work counter of the unit renderer
  int units_per_grab = 0;        // 0: automatic
  int unit_blocks = 0;           // 0: automatic
  bool autotune = false;         // tune renderPoseBatch's launch shape per workload
  float label_error = 0.01f;     // error the sampler network's labels were made for
  int sampler = kSamplerNetwork; // source of the adaptive renders' required samples
  double adaptive_ms[4] = {0, 0, 0, 0};  // stage times of the last adaptive render
  LaunchTuner tuner;
  std::string tune_cache;        // file the tuner's entries are saved to (empty: none)
  std::string name_buf;
};

State& st() {
  static State s;
  return s;
}

void set_error(const std::string& e) {
  st().last_error = e;
  std::cerr << "[PyEye] ERROR: " << e << std::endl;
}

void note(const std::string& msg) {
  if (st().verbose) std::cout << "[PyEye] " << msg << std::endl;
}

// Throws std::runtime_error on HIP failure; entry points catch it.
void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void ensure_hip() {
  State& s = st();
  if (s.hip_ready) return;
  int count = 0;
  const hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) throw std::runtime_error("no HIP device available");
  hip_check(hipSetDevice(0), "hipSetDevice");
  std::string serr;
  if (!no_alloc_scratch_init(0, &serr)) throw std::runtime_error(serr);
  hip_check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipMalloc(&s.d_pose, sizeof(Pose)), "hipMalloc pose");
  hip_check(hipMalloc(&s.d_stats, sizeof(unsigned long long) * 4), "hipMalloc stats");
  hip_check(hipMalloc(&s.d_queue, sizeof(uint32_t) * kQueueWords), "hipMalloc queue");
  hip_check(hipMalloc(&s.d_packet_counts, sizeof(unsigned long long) * 2), "hipMalloc packet counts");
  hip_check(hipMemset(s.d_packet_counts, 0, sizeof(unsigned long long) * 2), "clear packet counts");
  for (hipEvent_t& e : s.ev) hip_check(hipEventCreate(&e), "hipEventCreate");
  hip_check(hipHostMalloc(&s.h_fault, sizeof(uint32_t), hipHostMallocMapped | hipHostMallocCoherent),
            "hipHostMalloc fault word");
  *s.h_fault = 0u;
  hip_check(hipHostGetDevicePointer((void**)&s.d_fault, s.h_fault, 0), "fault word device address");
  s.hip_ready = true;
}

// The packet traversal flags a launch that entered it with a partial wave
// and returns misses there: the call that rendered fails instead of handing
// back those colours.  Called after the call's stream synchronisation.
void check_packet_fault(const char* what) {
  State& s = st();
  if (s.h_fault && __atomic_load_n(s.h_fault, __ATOMIC_ACQUIRE) != 0u) {
    *s.h_fault = 0u;
    throw std::runtime_error(std::string(what) +
                             ": packet traversal entered by a partial wave; the results were discarded");
  }
}

template <typename T>
void ensure_device(T** p, size_t* cap, size_t n) {
  if (*cap >= n && *p) return;
  if (*p) hip_check(hipFree(*p), "hipFree");
  *p = nullptr;
  *cap = 0;
  if (n == 0) return;
  hip_check(hipMalloc(p, sizeof(T) * n), "hipMalloc");
  *cap = n;
}

Cam* current_cam() {
  State& s = st();
  if (s.cams.empty()) return nullptr;
  return &s.cams[s.current];
}

EyeState* current_eye() {
  Cam* c = current_cam();
  if (!c || c->eye < 0) return nullptr;
  return &st().eyes[(size_t)c->eye];
}

void free_scene() {
  State& s = st();
  free_device_scene(&s.scene);
  if (s.d_pos) (void)hipFree(s.d_pos);
  if (s.d_col) (void)hipFree(s.d_col);
  if (s.d_mode) (void)hipFree(s.d_mode);
  if (s.d_texels) (void)hipFree(s.d_texels);
  if (s.d_textures) (void)hipFree(s.d_textures);
  s.d_pos = s.d_col = nullptr;
  s.d_mode = nullptr;
  s.d_texels = nullptr;
  s.d_textures = nullptr;
  for (EyeState& e : s.eyes) {
    if (e.d_omms) (void)hipFree(e.d_omms);
    if (e.d_offsets) (void)hipFree(e.d_offsets);
    if (e.d_samples) (void)hipFree(e.d_samples);
    if (e.d_colours) (void)hipFree(e.d_colours);
    if (e.d_proj_ids) (void)hipFree(e.d_proj_ids);
    if (e.d_chunk_partial) (void)hipFree(e.d_chunk_partial);
    if (e.d_chunk_done) (void)hipFree(e.d_chunk_done);
    if (e.d_split) (void)hipFree(e.d_split);
    if (e.d_counts) (void)hipFree(e.d_counts);
    if (e.d_sample_offsets) (void)hipFree(e.d_sample_offsets);
    if (e.d_order) (void)hipFree(e.d_order);
    if (e.d_nbr) (void)hipFree(e.d_nbr);
  }
  s.eyes.clear();
  s.cams.clear();
  s.current = 0;
  s.loaded = false;
}

// Long-task split of the packet task kernel (render.h, UnitLaunch::split_cost):
// 0 never, 1 always, 2 (default) chosen per eye by measurement.  Each trial
// traces renders 1 and 2 of the same tasks without it, render 3 records the
// tasks' steps, renders 4 and 5 trace with it; the faster way (by the
// shorter of each pair's trace times, with it only if 3% faster) is kept
// until render kSplitTrialEvery, when a new trial begins.  Any change of the
// rays (ommatidia, rate, scene) starts a trial over.  Results do not depend
// on it.
int g_task_split = 2;
constexpr uint32_t kSplitTrialEvery = 1000;

void split_restart(EyeState& e) {
  e.split_valid = false;
  e.split_frames = 0;
  e.split_ms[0] = e.split_ms[1] = 1e30f;
}

// Whether this render of e records its tasks' steps, and whether it also
// plans the split from the last render's steps.
void split_choice(const EyeState& e, bool* record, bool* plan) {
  bool on = false, rec = false;
  if (g_task_split == 1) {
    on = rec = true;
  } else if (g_task_split == 2) {
    const uint32_t k = e.split_frames;
    rec = k >= 3 && (k <= 5 || e.split_ms[1] < 0.97f * e.split_ms[0]);
    on = rec && k != 3;
  }
  *record = rec;
  *plan = on && e.split_valid;
}

// After a render through the task kernel: its trace time goes to the trial.
void split_account(EyeState& e, float trace_ms) {
  const uint32_t k = e.split_frames;
  if (k == 1 || k == 2) e.split_ms[0] = std::min(e.split_ms[0], trace_ms);
  if (k == 4 || k == 5) e.split_ms[1] = std::min(e.split_ms[1], trace_ms);
  if (k + 1 >= kSplitTrialEvery) {
    e.split_frames = 1;   // next trial: the buffers' costs stay valid
    e.split_ms[0] = e.split_ms[1] = 1e30f;
  } else {
    e.split_frames = k + 1;
  }
}

// Uploads what changed since the last render: the ommatidia and their
// processing order (host sort) only when the ommatidia changed, the count
// and sample-record offsets when the sample counts (or the uniform rate)
// changed.
void upload_eye(EyeState& e) {
  State& s = st();
  if (!e.device_dirty()) return;
  const size_t n = e.omms.size();
  std::vector<uint32_t> order;
  if (e.omm_dirty) {
    ensure_device(&e.d_omms, &e.d_omms_cap, n);
    if (n)
      hip_check(hipMemcpyAsync(e.d_omms, e.omms.data(), sizeof(Ommatidium) * n, hipMemcpyHostToDevice, s.stream),
                "upload ommatidia");
    // Units are processed in the order of (position cell, axis Morton code):
    // waves running at the same time trace neighbouring view directions from
    // the same place and share BVH nodes in cache.  Cells are 1 scene unit (or
    // 1/1024 of the positions' extent, if larger), so one eye (millimetres
    // across) sorts by direction alone while eyes merged into one ommatidium
    // list stay grouped.  Output positions are unchanged.
    float plo[3] = {INFINITY, INFINITY, INFINITY}, phi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < n; ++i) {
      const float p[3] = {e.omms[i].px, e.omms[i].py, e.omms[i].pz};
      for (int a = 0; a < 3; ++a)
        if (std::isfinite(p[a])) {
          plo[a] = std::min(plo[a], p[a]);
          phi[a] = std::max(phi[a], p[a]);
        }
    }
    double ext = 0.0;
    for (int a = 0; a < 3; ++a)
      if (phi[a] >= plo[a]) ext = std::max(ext, (double)phi[a] - (double)plo[a]);
    const double cell = std::max(1.0, ext / 1023.0);
    auto cell_of = [&](float p, int a) -> uint32_t {
      if (!std::isfinite(p) || !(phi[a] >= plo[a])) return 0u;
      const double q = std::floor(((double)p - (double)plo[a]) / cell);
      return (uint32_t)std::min(1023.0, std::max(0.0, q));
    };
    std::vector<std::pair<uint64_t, uint32_t>> key(n);
    for (size_t i = 0; i < n; ++i) {
      const V3 d = normalize(v3(e.omms[i].dx, e.omms[i].dy, e.omms[i].dz));
      const float lo[3] = {-1.0f, -1.0f, -1.0f}, hi[3] = {1.0f, 1.0f, 1.0f};
      const bool ok = d.x == d.x && d.y == d.y && d.z == d.z;
      const uint64_t where = ((uint64_t)expand_bits10(cell_of(e.omms[i].px, 0)) << 2) |
                             ((uint64_t)expand_bits10(cell_of(e.omms[i].py, 1)) << 1) |
                             (uint64_t)expand_bits10(cell_of(e.omms[i].pz, 2));
      key[i] = {(where << 30) | (ok ? morton30(d.x, d.y, d.z, lo, hi) : 0u), (uint32_t)i};
    }
    std::sort(key.begin(), key.end());
    order.resize(n);
    for (size_t i = 0; i < n; ++i) order[i] = key[i].second;
    ensure_device(&e.d_order, &e.d_order_cap, n);
    if (n)
      hip_check(hipMemcpyAsync(e.d_order, order.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s.stream),
                "upload order");
  }
  std::vector<uint64_t> first;
  if (e.counts_dirty || e.omm_dirty) {
    first.resize(n);
    if (!e.counts.empty()) {
      ensure_device(&e.d_offsets, &e.d_offsets_cap, e.offsets.size());
      hip_check(hipMemcpyAsync(e.d_offsets, e.offsets.data(), sizeof(uint32_t) * e.offsets.size(),
                               hipMemcpyHostToDevice, s.stream),
                "upload offsets");
      ensure_device(&e.d_counts, &e.d_counts_cap, n);
      hip_check(hipMemcpyAsync(e.d_counts, e.counts.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s.stream),
                "upload counts");
      for (size_t i = 0; i < n; ++i) first[i] = e.offsets[i];
    } else {
      for (size_t i = 0; i < n; ++i) first[i] = (uint64_t)i * (uint64_t)e.spp;
    }
    ensure_device(&e.d_sample_offsets, &e.d_sample_offsets_cap, n);
    if (n)
      hip_check(hipMemcpyAsync(e.d_sample_offsets, first.data(), sizeof(uint64_t) * n, hipMemcpyHostToDevice,
                               s.stream),
                "upload sample offsets");
  }
  hip_check(hipStreamSynchronize(s.stream), "sync");
  e.omm_dirty = false;
  e.counts_dirty = false;
  split_restart(e);   // other rays: the last render's task costs no longer apply
}

EyeLaunch launch_of(EyeState& e, const Pose* d_poses, int pose_count) {
  EyeLaunch L;
  L.omms = e.d_omms;
  L.omm_count = (int)e.omms.size();
  L.offsets = e.counts.empty() ? nullptr : e.d_offsets;
  L.spp = e.spp;
  L.rays_per_pose = e.total_samples();
  L.poses = d_poses;
  L.pose_count = pose_count;
  L.pose_base = 0;
  L.seed = st().seed;
  L.eye = e.eye_index;
  L.frame = e.frame;
  return L;
}

UnitLaunch unit_launch_of(EyeState& e, const Pose* d_poses, int pose_count) {
  State& s = st();
  UnitLaunch U;
  U.omms = e.d_omms;
  U.omm_count = (int)e.omms.size();
  U.poses = d_poses;
  U.pose_count = pose_count;
  U.seed = s.seed;
  U.eye = e.eye_index;
  U.frame = e.frame;
  U.sample_base = e.sample_base;
  U.spp = e.spp;
  U.omm_counts = e.counts.empty() ? nullptr : e.d_counts;
  U.colours = e.d_colours;
  U.queue = s.d_queue;
  U.units_per_grab = s.units_per_grab;
  U.blocks = s.unit_blocks;
  U.traversal = unit_traversal_of(trace_mode());
  U.order = e.d_order;
  U.packet_counts = s.count_packets ? s.d_packet_counts : nullptr;
  return U;
}

void ensure_frame_buffer() {
  State& s = st();
  const size_t n = (size_t)std::max(s.width, 0) * (size_t)std::max(s.height, 0);
  if (s.d_frame_cap < n || !s.d_frame) {
    ensure_device(&s.d_frame, &s.d_frame_cap, n);
    if (n) hip_check(hipMemsetAsync(s.d_frame, 0, sizeof(uchar4) * n, s.stream), "clear frame");
  }
}

void build_scene_bvh() {
  State& s = st();
  std::string err;
  hip_check(hipStreamSynchronize(s.stream), "sync");
  if (!build_device_scene(s.d_pos, s.d_col, s.d_mode, (int)s.tri_count, s.leaf_max, s.stream, &s.scene,
                          &err))
    throw std::runtime_error(err);
  static uint64_t generations = 0;
  s.scene.generation = ++generations;
  s.scene.fault = s.d_fault;
  s.scene.sky_mode = s.host.sky_mode;
  s.scene.texels = s.d_texels;
  s.scene.textures = s.d_textures;
  s.scene.texture_count = (int)s.host.textures.size();
}

double render_compound(EyeState& e, const Cam& c) {
  State& s = st();
  upload_eye(e);
  const uint64_t total = e.total_samples();
  // per-sample records are kept only when asked for (or for the raw projection)
  const bool record = s.record_samples || e.shader == "raw_ommatidial_samples";
  if (record) ensure_device(&e.d_samples, &e.d_samples_cap, (size_t)total);
  ensure_device(&e.d_colours, &e.d_colours_cap, e.omms.size());
  ensure_frame_buffer();
  hip_check(hipMemcpyAsync(s.d_pose, &c.pose, sizeof(Pose), hipMemcpyHostToDevice, s.stream), "upload pose");
  EyeLaunch L = launch_of(e, s.d_pose, 1);
  std::string err;
  const auto t0 = std::chrono::steady_clock::now();
  hip_check(hipEventRecord(s.ev[0], s.stream), "event");
  UnitLaunch U = unit_launch_of(e, s.d_pose, 1);
  if (record) {
    U.samples = e.d_samples;
    U.sample_offsets = e.d_sample_offsets;
  }
  if (e.counts.empty()) {
    // one pose: let the chunks of each ommatidium's samples spread over waves
    const size_t chunks = ((size_t)e.spp + 63) / 64;
    ensure_device(&e.d_chunk_partial, &e.d_chunk_partial_cap, e.omms.size() * chunks * 64);
    ensure_device(&e.d_chunk_done, &e.d_chunk_done_cap, e.omms.size());
    U.chunk_partial = e.d_chunk_partial;
    U.chunk_partial_cap = e.d_chunk_partial_cap;
    U.chunk_done = e.d_chunk_done;
    U.chunk_done_cap = e.d_chunk_done_cap;
    // the packet task kernel's long-task split, planned from the last render
    const uint64_t tasks = (uint64_t)e.omms.size() * chunks;
    if (e.split_tasks != tasks || e.split_generation != s.scene.generation) {
      split_restart(e);
      e.split_tasks = tasks;
      e.split_generation = s.scene.generation;
    }
    bool rec = false, plan = false;
    split_choice(e, &rec, &plan);
    if (rec) {
      const uint32_t* before = e.d_split;
      ensure_device(&e.d_split, &e.d_split_cap, (size_t)split_buffer_words(tasks));
      if (e.d_split != before) {   // new buffer: no steps recorded in it yet
        e.split_valid = false;
        plan = false;
      }
      U.split_cost = e.d_split;
      U.split_flag = U.split_cost + tasks;
      U.split_list = U.split_flag + tasks;
      U.split_count = U.split_list + kSplitListMax;
      U.split_sub_done = U.split_count + kSplitCountWords;
      U.split_tasks = tasks;
      U.split_plan = plan;
    }
  }
  if (!launch_render_units(s.scene, U, s.stream, &err)) {
    split_restart(e);
    throw std::runtime_error(err);
  }
  e.split_listed[0] = e.split_listed[1] = e.split_listed[2] = 0;
  const bool tasks_kernel = std::strcmp(last_render_kernel(), "k_render_tasks") == 0;
  if (U.split_cost && tasks_kernel) {
    if (U.split_plan)
      hip_check(hipMemcpyAsync(e.split_listed, U.split_count, sizeof(e.split_listed), hipMemcpyDeviceToHost,
                               s.stream),
                "split count");
    e.split_valid = true;
  } else {
    e.split_valid = false;
  }
  if (!tasks_kernel) split_restart(e);
  hip_check(hipEventRecord(s.ev[1], s.stream), "event");
  hip_check(hipEventRecord(s.ev[2], s.stream), "event");
  bool ok = true;
  int sph = -1;
  bool sph_ids = false;
  if (e.shader == "spherical_positionwise" || e.shader == "spherical_positionwise_ids") sph = kSphericalPosition;
  if (e.shader == "spherical_orientationwise" || e.shader == "spherical_orientationwise_ids")
    sph = kSphericalOrientation;
  if (e.shader == "spherical_split_orientationwise") sph = kSphericalSplit;
  if (sph >= 0) sph_ids = e.shader.size() > 4 && e.shader.compare(e.shader.size() - 4, 4, "_ids") == 0;
  if (sph >= 0) {
    const int pixels = s.width * s.height;
    if (e.proj_w != s.width || e.proj_h != s.height || e.proj_variant != sph || e.proj_version != e.omm_version) {
      ensure_device(&e.d_proj_ids, &e.d_proj_cap, (size_t)std::max(pixels, 1));
      ok = launch_spherical_ids(e.d_omms, L.omm_count, s.width, s.height, sph, e.d_proj_ids, s.stream, &err);
      e.proj_w = s.width;
      e.proj_h = s.height;
      e.proj_variant = sph;
      e.proj_version = e.omm_version;
    }
    if (ok) ok = launch_frame_lookup(e.d_colours, e.d_proj_ids, pixels, sph_ids, s.d_frame, s.stream, &err);
  } else if (e.shader == "single_dimension")
    ok = launch_frame_single_dim(e.d_colours, L.omm_count, s.d_frame, s.width, s.height, s.stream, &err);
  else if (e.shader == "raw_ommatidial_samples")
    ok = launch_frame_raw_samples(e.d_samples, e.counts.empty() ? nullptr : e.d_offsets, L.omm_count, e.spp,
                                  s.d_frame, s.width, s.height, s.stream, &err);
  else
    ok = launch_frame_single_dim_fast(e.d_colours, L.omm_count, s.d_frame, s.width, s.stream, &err);
  if (!ok) throw std::runtime_error(err);
  hip_check(hipEventRecord(s.ev[3], s.stream), "event");
  hip_check(hipStreamSynchronize(s.stream), "render");
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  check_packet_fault("renderFrame");
  float f[3] = {0, 0, 0};
  for (int k = 0; k < 3; ++k) (void)hipEventElapsedTime(&f[k], s.ev[k], s.ev[k + 1]);
  s.timings[0] = f[0];
  s.timings[1] = f[1];
  s.timings[2] = f[2];
  s.timings[3] = (double)f[0] + f[1] + f[2];
  if (e.counts.empty() && std::strcmp(last_render_kernel(), "k_render_tasks") == 0) split_account(e, f[0]);
  e.frame += 1;
  e.has_render = true;
  e.rendered_rays = record ? total : 0;
  s.frame_host_stale = true;
  return ms;
}

CameraLaunch camera_launch_of(const Cam& c) {
  State& s = st();
  CameraLaunch C;
  C.pose = c.pose;
  C.kind = c.kind == kOrthographic ? kCameraOrthographic : c.kind == kPanoramic ? kCameraPanoramic : kCameraPinhole;
  C.scale_x = c.scale[0];
  C.scale_y = c.scale[1];
  C.scale_z = c.scale[2];
  C.start_radius = c.start_radius;
  C.width = s.width;
  C.height = s.height;
  return C;
}

// Pinhole, orthographic and panoramic cameras: one primary ray per pixel.
double render_regular(const Cam& c) {
  State& s = st();
  ensure_frame_buffer();
  std::string err;
  const auto t0 = std::chrono::steady_clock::now();
  hip_check(hipEventRecord(s.ev[0], s.stream), "event");
  if (!launch_camera_frame(s.scene, camera_launch_of(c), s.d_frame, nullptr, s.stream, &err))
    throw std::runtime_error(err);
  hip_check(hipEventRecord(s.ev[1], s.stream), "event");
  hip_check(hipStreamSynchronize(s.stream), "render");
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  float f = 0.0f;
  (void)hipEventElapsedTime(&f, s.ev[0], s.ev[1]);
  s.timings[0] = f;
  s.timings[1] = 0.0;
  s.timings[2] = 0.0;
  s.timings[3] = f;
  s.frame_host_stale = true;
  return ms;
}

void recompute_offsets(EyeState& e) {
  e.offsets.clear();
  if (e.counts.empty()) return;
  e.offsets.resize(e.counts.size() + 1);
  uint64_t run = 0;
  for (size_t i = 0; i < e.counts.size(); ++i) {
    e.offsets[i] = (uint32_t)run;
    run += e.counts[i];
  }
  if (run > 0xffffffffull) throw std::runtime_error("more than 2^32 samples per eye");
  e.offsets.back() = (uint32_t)run;
}

void sync_frame_host() {
  State& s = st();
  const size_t n = (size_t)std::max(s.width, 0) * (size_t)std::max(s.height, 0);
  if (s.frame_host.size() != n * 4) s.frame_host.assign(n * 4, 0);
  if (!s.frame_host_stale || !s.hip_ready || !s.d_frame || n == 0) return;
  hip_check(hipMemcpy(s.frame_host.data(), s.d_frame, n * 4, hipMemcpyDeviceToHost), "frame readback");
  s.frame_host_stale = false;
}

// Scoped device copies of host arrays (freed on scope exit).
struct DeviceTemp {
  std::vector<void*> ptrs;
  ~DeviceTemp() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  T* upload(const T* host, size_t n) {
    T* d = nullptr;
    hip_check(hipMalloc(&d, sizeof(T) * (n ? n : 1)), "hipMalloc");
    ptrs.push_back(d);
    if (host && n) hip_check(hipMemcpy(d, host, sizeof(T) * n, hipMemcpyHostToDevice), "upload");
    return d;
  }
};

#define ABI_GUARD_BEGIN try {
#define ABI_GUARD_END(ret)                      \
  }                                             \
  catch (const std::exception& ex) {            \
    set_error(ex.what());                       \
    ret;                                        \
  }                                             \
  catch (...) {                                 \
    set_error("unknown error");                 \
    ret;                                        \
  }

}  // namespace

// The learned sampler's neighbour table on the device (kNeighbours per
// ommatidium), rebuilt on the host when the eye's ommatidia have changed.
static const uint32_t* neighbours_of(EyeState& e) {
  State& s = st();
  if (e.d_nbr && e.nbr_version == e.omm_version) return e.d_nbr;
  const size_t n = e.omms.size();
  std::vector<uint32_t> t(std::max<size_t>(n, 1) * kNeighbours, 0u);
  nearest_ommatidia(reinterpret_cast<const float*>(e.omms.data()), n, kNeighbours, kNeighbourPosScale, t.data());
  ensure_device(&e.d_nbr, &e.d_nbr_cap, t.size());
  hip_check(hipMemcpyAsync(e.d_nbr, t.data(), sizeof(uint32_t) * t.size(), hipMemcpyHostToDevice, s.stream),
            "upload neighbours");
  hip_check(hipStreamSynchronize(s.stream), "sync");
  e.nbr_version = e.omm_version;
  return e.d_nbr;
}

extern "C" {

void setVerbosity(bool v) { st().verbose = v; }

void loadGlTFscene(const char* filepath) {
  ABI_GUARD_BEGIN
  State& s = st();
  if (!filepath) throw std::runtime_error("loadGlTFscene: null path");
  HostScene hs;
  std::string err;
  if (!load_gltf(filepath, &hs, &err)) throw std::runtime_error(err);
  for (const std::string& w : hs.warnings) std::cerr << "[PyEye] glTF warning: " << w << std::endl;
  ensure_hip();
  free_scene();
  s.host = std::move(hs);
  const size_t n = s.host.tri_count();
  s.tri_count = n;
  if (n) {
    hip_check(hipMalloc(&s.d_pos, sizeof(float) * 9 * n), "hipMalloc positions");
    hip_check(hipMalloc(&s.d_col, sizeof(float) * 9 * n), "hipMalloc colours");
    hip_check(hipMalloc(&s.d_mode, sizeof(uint32_t) * n), "hipMalloc modes");
    hip_check(hipMemcpy(s.d_pos, s.host.pos.data(), sizeof(float) * 9 * n, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(s.d_col, s.host.col.data(), sizeof(float) * 9 * n, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(s.d_mode, s.host.mode.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice), "upload");
  }
  if (!s.host.textures.empty()) {
    std::vector<TexView> tv;
    for (const TextureDesc& d : s.host.textures)
      tv.push_back(TexView{d.width, d.height, d.wrap_s, d.wrap_t, d.filter, d.first});
    hip_check(hipMalloc(&s.d_texels, sizeof(uint32_t) * s.host.texels.size()), "hipMalloc texels");
    hip_check(hipMalloc(&s.d_textures, sizeof(TexView) * tv.size()), "hipMalloc textures");
    hip_check(hipMemcpy(s.d_texels, s.host.texels.data(), sizeof(uint32_t) * s.host.texels.size(),
                        hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(s.d_textures, tv.data(), sizeof(TexView) * tv.size(), hipMemcpyHostToDevice), "upload");
    s.host.texels.clear();
    s.host.texels.shrink_to_fit();
  }
  build_scene_bvh();
  // release the host copies of the geometry; the device keeps them for rebuilds
  s.host.pos.clear(); s.host.pos.shrink_to_fit();
  s.host.col.clear(); s.host.col.shrink_to_fit();
  s.host.mode.clear(); s.host.mode.shrink_to_fit();
  uint32_t eye_idx = 0;
  for (CameraDesc& cd : s.host.cameras) {
    Cam c;
    c.name = cd.name;
    c.kind = cd.kind;
    c.pose = cd.pose;
    c.yfov_deg = cd.yfov_deg;
    if (cd.kind == kPerspective) {
      // setYFOV (PerspectiveCamera.cpp:19-24) with aspect ratio 1
      const float yfov = cd.yfov_deg / 180.0f * kPiF;
      c.scale[1] = std::tan(yfov / 2.0f) * c.scale[2];
      c.scale[0] = c.scale[1] * 1.0f;
    } else if (cd.kind == kOrthographic) {
      c.scale[0] = cd.xmag;
      c.scale[1] = cd.ymag;
    }
    if (cd.kind == kCompound) {
      EyeState e;
      e.omms = std::move(cd.ommatidia);
      e.eye_index = eye_idx++;
      e.data_path = cd.eye_path;
      e.shader = cd.projection;
      c.eye = (int)s.eyes.size();
      s.eyes.push_back(std::move(e));
    }
    s.cams.push_back(std::move(c));
  }
  s.host.cameras.clear();
  if (s.cams.empty()) {  // MulticamScene::getCamera adds a default perspective camera
    Cam c;
    c.name = "Default Camera";
    pose_reset(&c.pose);
    s.cams.push_back(c);
  }
  s.current = 0;
  s.loaded = true;
  s.last_error.clear();
  note("Loaded scene '" + std::string(filepath) + "': " + std::to_string(n) + " triangles, " +
       std::to_string(s.cams.size()) + " cameras, " + std::to_string(s.scene.node_count) + " BVH nodes.");
  ABI_GUARD_END(return)
}

void stop(void) {
  ABI_GUARD_BEGIN
  note("Cleaning eye renderer resources.");
  State& s = st();
  if (s.hip_ready) {
    (void)hipStreamSynchronize(s.stream);
    free_scene();
    s.batch.release();
    if (s.d_frame) (void)hipFree(s.d_frame);
    s.d_frame = nullptr;
    s.d_frame_cap = 0;
  }
  ABI_GUARD_END(return)
}

void setRenderSize(int w, int h) {
  ABI_GUARD_BEGIN
  State& s = st();
  s.width = std::max(w, 0);
  s.height = std::max(h, 0);
  note("Resizing rendering buffer to (" + std::to_string(w) + ", " + std::to_string(h) + ").");
  s.frame_host.assign((size_t)s.width * (size_t)s.height * 4, 0);
  if (s.hip_ready) {
    if (s.d_frame) hip_check(hipFree(s.d_frame), "hipFree");
    s.d_frame = nullptr;
    s.d_frame_cap = 0;
    ensure_frame_buffer();
  }
  s.frame_host_stale = false;
  ABI_GUARD_END(return)
}

double renderFrame(void) {
  ABI_GUARD_BEGIN
  State& s = st();
  if (!s.loaded) throw std::runtime_error("renderFrame: no scene loaded");
  Cam* c = current_cam();
  double ms = 0.0;
  if (c && c->eye >= 0) {
    ms = render_compound(s.eyes[(size_t)c->eye], *c);
  } else if (c) {
    ms = render_regular(*c);
  }
  s.last_error.clear();
  if (s.verbose) std::cout << "[PyEye] Rendered frame in " << ms << "ms." << std::endl;
  return ms;
  ABI_GUARD_END(return -1.0)
}

void displayFrame(void) {}

void saveFrameAs(char* ppmFilename) {
  ABI_GUARD_BEGIN
  State& s = st();
  if (!ppmFilename) throw std::runtime_error("saveFrameAs: null file name");
  sync_frame_host();
  FILE* f = fopen(ppmFilename, "wb");
  if (!f) throw std::runtime_error(std::string("saveFrameAs: cannot open '") + ppmFilename + "'");
  fprintf(f, "P6\n%d %d\n255\n", s.width, s.height);
  // rows are written bottom-up (the frame is stored upside down, sutil.cpp:82-103)
  std::vector<uint8_t> row((size_t)s.width * 3);
  for (int y = s.height - 1; y >= 0; --y) {
    for (int x = 0; x < s.width; ++x)
      for (int k = 0; k < 3; ++k) row[(size_t)x * 3 + k] = s.frame_host[((size_t)y * s.width + x) * 4 + k];
    fwrite(row.data(), 1, row.size(), f);
  }
  fclose(f);
  note(std::string("Saved render as '") + ppmFilename + "'");
  ABI_GUARD_END(return)
}

unsigned char* getFramePointer(void) {
  ABI_GUARD_BEGIN
  note("Retrieving frame pointer...");
  sync_frame_host();
  State& s = st();
  return s.frame_host.empty() ? nullptr : s.frame_host.data();
  ABI_GUARD_END(return nullptr)
}

size_t getCameraCount(void) { return st().cams.size(); }

void nextCamera(void) { gotoCamera((int)st().current + 1); }
void previousCamera(void) { gotoCamera((int)st().current - 1); }
size_t getCurrentCameraIndex(void) { return st().current; }

const char* getCurrentCameraName(void) {
  Cam* c = current_cam();
  return c ? c->name.c_str() : "";
}

void gotoCamera(int index) {
  State& s = st();
  const int n = (int)s.cams.size();
  if (n == 0) return;
  s.current = (size_t)(((index % n) + n) % n);
}

bool gotoCameraByName(char* name) {
  State& s = st();
  if (!name) return false;
  for (size_t i = 0; i < s.cams.size(); ++i)
    if (s.cams[i].name == name) {
      s.current = i;
      return true;
    }
  // the reference leaves the camera index past the last camera, i.e. at 0
  if (!s.cams.empty()) s.current = 0;
  return false;
}

void setCameraPosition(float x, float y, float z) {
  Cam* c = current_cam();
  if (!c) return;
  c->pose.pos[0] = x; c->pose.pos[1] = y; c->pose.pos[2] = z;
}

void getCameraPosition(float* x, float* y, float* z) {
  Cam* c = current_cam();
  const float p[3] = {c ? c->pose.pos[0] : 0.0f, c ? c->pose.pos[1] : 0.0f, c ? c->pose.pos[2] : 0.0f};
  if (x) *x = p[0];
  if (y) *y = p[1];
  if (z) *z = p[2];
}

void setCameraLocalSpace(float lxx, float lxy, float lxz, float lyx, float lyy, float lyz, float lzx, float lzy,
                         float lzz) {
  Cam* c = current_cam();
  if (!c) return;
  c->pose.x[0] = lxx; c->pose.x[1] = lxy; c->pose.x[2] = lxz;
  c->pose.y[0] = lyx; c->pose.y[1] = lyy; c->pose.y[2] = lyz;
  c->pose.z[0] = lzx; c->pose.z[1] = lzy; c->pose.z[2] = lzz;
}

void rotateCameraAround(float angle, float x, float y, float z) {
  Cam* c = current_cam();
  if (c) pose_rotate_around(&c->pose, angle, x, y, z);
}

void rotateCameraLocallyAround(float angle, float x, float y, float z) {
  Cam* c = current_cam();
  if (!c) return;
  // transformToLocal, DataRecordCamera.h:57-65
  const Pose& p = c->pose;
  const V3 a = (v3(p.x[0], p.x[1], p.x[2]) * x + v3(p.y[0], p.y[1], p.y[2]) * y) + v3(p.z[0], p.z[1], p.z[2]) * z;
  pose_rotate_around(&c->pose, angle, a.x, a.y, a.z);
}

void translateCamera(float x, float y, float z) {
  Cam* c = current_cam();
  if (c) pose_move(&c->pose, x, y, z);
}

void translateCameraLocally(float x, float y, float z) {
  Cam* c = current_cam();
  if (!c) return;
  const Pose& p = c->pose;
  const V3 a = (v3(p.x[0], p.x[1], p.x[2]) * x + v3(p.y[0], p.y[1], p.y[2]) * y) + v3(p.z[0], p.z[1], p.z[2]) * z;
  pose_move(&c->pose, a.x, a.y, a.z);
}

void resetCameraPose(void) {
  Cam* c = current_cam();
  if (c) pose_reset(&c->pose);
}

void setCameraPose(float posX, float posY, float posZ, float rotX, float rotY, float rotZ) {
  Cam* c = current_cam();
  if (!c) return;
  pose_reset(&c->pose);
  pose_rotate_around(&c->pose, rotX, 1.0f, 0.0f, 0.0f);
  pose_rotate_around(&c->pose, rotY, 0.0f, 1.0f, 0.0f);
  pose_rotate_around(&c->pose, rotZ, 0.0f, 0.0f, 1.0f);
  pose_move(&c->pose, posX, posY, posZ);
}

bool isCompoundEyeActive(void) { return current_eye() != nullptr; }

void setCurrentEyeSamplesPerOmmatidium(int s) {
  EyeState* e = current_eye();
  if (!e) return;
  const int spp = std::max(1, s);
  // per-sample record offsets depend on the rate (with uniform counts)
  if (spp != e->spp && e->counts.empty()) e->counts_dirty = true;
  e->spp = spp;
  e->restart_streams();
}

int getCurrentEyeSamplesPerOmmatidium(void) {
  EyeState* e = current_eye();
  return e ? e->spp : -1;
}

void changeCurrentEyeSamplesPerOmmatidiumBy(int s) {
  EyeState* e = current_eye();
  if (!e) return;
  setCurrentEyeSamplesPerOmmatidium(e->spp + s);
}

size_t getCurrentEyeOmmatidialCount(void) {
  EyeState* e = current_eye();
  return e ? e->omms.size() : 0;
}

void setOmmatidia(OmmatidiumPacket* omms, size_t count) {
  ABI_GUARD_BEGIN
  EyeState* e = current_eye();
  if (!e) return;
  if (count && !omms) throw std::runtime_error("setOmmatidia: null array");
  static_assert(sizeof(OmmatidiumPacket) == sizeof(Ommatidium), "packet layout");
  if (count != e->omms.size()) {
    e->restart_streams();  // CompoundEye::reconfigureOmmatidialCount reallocates the random states
    e->counts.clear();
    e->offsets.clear();
  }
  e->omms.resize(count);
  e->omm_version++;
  if (count) memcpy(e->omms.data(), omms, sizeof(Ommatidium) * count);
  e->omm_dirty = true;
  ABI_GUARD_END(return)
}

const char* getCurrentEyeDataPath(void) {
  EyeState* e = current_eye();
  return e ? e->data_path.c_str() : "";
}

void setCurrentEyeShaderName(char* name) {
  EyeState* e = current_eye();
  if (!e || !name) return;
  e->shader = name;
}

bool isInsideHitGeometry(float x, float y, float z, char* name) {
  ABI_GUARD_BEGIN
  State& s = st();
  if (!name) return false;
  const HitboxMesh* hb = nullptr;
  for (const HitboxMesh& h : s.host.hitboxes)
    if (h.name == name) { hb = &h; break; }
  if (!hb) {
    std::cerr << "WARNING: No hitbox with the given name \"" << name << "\" is present in the scene." << std::endl;
    return false;
  }
  // Parity of crossings of a +x ray from outside the mesh bounds to the point
  // (sutil/hitscanprocessing.cpp:20-82), evaluated in world space.
  const V3 p = v3(x, y, z);
  const V3 start = v3(hb->lo[0] - 1.0f, y, z);
  const V3 dir = normalize(p - start);
  unsigned crossings = 0;
  for (size_t t = 0; t + 9 <= hb->tris.size(); t += 9) {
    const V3 p0 = v3(hb->tris[t], hb->tris[t + 1], hb->tris[t + 2]);
    const V3 p1 = v3(hb->tris[t + 3], hb->tris[t + 4], hb->tris[t + 5]);
    const V3 p2 = v3(hb->tris[t + 6], hb->tris[t + 7], hb->tris[t + 8]);
    const V3 nrm = normalize(cross(p1 - p0, p2 - p0));
    const float den = dot(nrm, dir);
    if (den == 0.0f) continue;
    const float dist = dot(p0 - start, nrm) / den;
    if (dist == 0.0f) continue;
    const V3 hit = start + dir * dist;
    if (dist < 0.0f || hit.x > p.x) continue;
    if (dot(nrm, cross(p1 - p0, hit - p0)) < 0.0f) continue;
    if (dot(nrm, cross(p2 - p1, hit - p1)) < 0.0f) continue;
    if (dot(nrm, cross(p0 - p2, hit - p2)) < 0.0f) continue;
    ++crossings;
  }
  return (crossings % 2u) == 1u;
  ABI_GUARD_END(return false)
}

static EyeFloat3 bounds_of(const char* name, bool max_side) {
  State& s = st();
  EyeFloat3 r = {0.0f, 0.0f, 0.0f};
  if (!name) return r;
  for (const HitboxMesh& h : s.host.hitboxes)
    if (h.name == name) {
      const float* b = max_side ? h.hi : h.lo;
      r.x = b[0]; r.y = b[1]; r.z = b[2];
      return r;
    }
  for (size_t m = 0; m < s.host.mesh_names.size(); ++m)
    if (s.host.mesh_names[m] == name && m < s.host.mesh_bounds.size()) {
      const float* b = s.host.mesh_bounds[m].data() + (max_side ? 3 : 0);
      r.x = b[0]; r.y = b[1]; r.z = b[2];
      return r;
    }
  return r;
}

EyeFloat3 getGeometryMaxBounds(char* name) { return bounds_of(name, true); }
EyeFloat3 getGeometryMinBounds(char* name) { return bounds_of(name, false); }

void* getWindowPointer(void) { return nullptr; }

const char* eyeRendererGetLastError(void) { return st().last_error.c_str(); }

size_t getOmmatidialColours(float* rgb, size_t count) {
  ABI_GUARD_BEGIN
  EyeState* e = current_eye();
  if (!e || !e->has_render || !rgb) return 0;
  const size_t n = std::min(count, e->omms.size());
  std::vector<float4> tmp(n);
  if (n) hip_check(hipMemcpy(tmp.data(), e->d_colours, sizeof(float4) * n, hipMemcpyDeviceToHost), "readback");
  for (size_t i = 0; i < n; ++i) {
    rgb[3 * i] = tmp[i].x;
    rgb[3 * i + 1] = tmp[i].y;
    rgb[3 * i + 2] = tmp[i].z;
  }
  return n;
  ABI_GUARD_END(return 0)
}

void setCurrentEyeSampleCounts(const uint32_t* counts, size_t count) {
  ABI_GUARD_BEGIN
  EyeState* e = current_eye();
  if (!e) return;
  if (count == 0) {
    e->counts.clear();
    e->offsets.clear();
  } else {
    if (count != e->omms.size())
      throw std::runtime_error("setCurrentEyeSampleCounts: count must equal the ommatidial count");
    if (!counts) throw std::runtime_error("setCurrentEyeSampleCounts: null array");
    e->counts.assign(counts, counts + count);
    recompute_offsets(*e);
  }
  e->restart_streams();
  e->counts_dirty = true;
  ABI_GUARD_END(return)
}

uint64_t getCurrentEyeTotalSamples(void) {
  EyeState* e = current_eye();
  return e ? e->total_samples() : 0;
}

void setCurrentEyeSampleBase(uint32_t base) {
  EyeState* e = current_eye();
  if (e) e->sample_base = base;
}

void setRandomSeed(uint32_t seed) { st().seed = seed; }

uint32_t getCurrentEyeFrameIndex(void) {
  EyeState* e = current_eye();
  return e ? e->frame : 0;
}

void setCurrentEyeFrameIndex(uint32_t frame) {
  EyeState* e = current_eye();
  if (e) e->frame = frame;
}

uint32_t getCurrentEyeIndex(void) {
  EyeState* e = current_eye();
  return e ? e->eye_index : 0xffffffffu;
}

int getLastRenderTimings(double* out4) {
  if (!out4) return -1;
  for (int k = 0; k < 4; ++k) out4[k] = st().timings[k];
  return 0;
}

int getCurrentEyeTraversalStats(uint64_t* out3) {
  ABI_GUARD_BEGIN
  State& s = st();
  EyeState* e = current_eye();
  Cam* c = current_cam();
  if (!e || !out3 || !s.loaded) return -1;
  upload_eye(*e);
  hip_check(hipMemsetAsync(s.d_stats, 0, sizeof(unsigned long long) * 4, s.stream), "clear stats");
  hip_check(hipMemcpyAsync(s.d_pose, &c->pose, sizeof(Pose), hipMemcpyHostToDevice, s.stream), "upload pose");
  EyeLaunch L = launch_of(*e, s.d_pose, 1);
  std::string err;
  if (!launch_trace_stats(s.scene, L, s.d_stats, s.stream, &err)) throw std::runtime_error(err);
  unsigned long long h[4];
  hip_check(hipMemcpyAsync(h, s.d_stats, sizeof(h), hipMemcpyDeviceToHost, s.stream), "stats readback");
  hip_check(hipStreamSynchronize(s.stream), "stats");
  out3[0] = h[0];
  out3[1] = h[1];
  out3[2] = h[2];
  return 0;
  ABI_GUARD_END(return -1)
}

int getCameraPose(float* out12) {
  Cam* c = current_cam();
  if (!c || !out12) return -1;
  static_assert(sizeof(Pose) == 12 * sizeof(float), "pose layout");
  memcpy(out12, &c->pose, sizeof(Pose));
  return 0;
}

size_t getCurrentEyeRays(float* rays, size_t count) {
  ABI_GUARD_BEGIN
  State& s = st();
  EyeState* e = current_eye();
  Cam* c = current_cam();
  if (!e || !rays || !s.loaded) return 0;
  upload_eye(*e);
  const uint64_t total = e->total_samples();
  const size_t n = (size_t)std::min<uint64_t>(total, count);
  if (n == 0) return 0;
  Ray* d_rays = nullptr;
  hip_check(hipMalloc(&d_rays, sizeof(Ray) * total), "hipMalloc rays");
  hip_check(hipMemcpyAsync(s.d_pose, &c->pose, sizeof(Pose), hipMemcpyHostToDevice, s.stream), "upload pose");
  EyeLaunch L = launch_of(*e, s.d_pose, 1);
  std::string err;
  const bool ok = launch_generate_rays(L, d_rays, s.stream, &err);
  hipError_t he = hipStreamSynchronize(s.stream);
  if (ok && he == hipSuccess) he = hipMemcpy(rays, d_rays, sizeof(Ray) * n, hipMemcpyDeviceToHost);
  (void)hipFree(d_rays);
  if (!ok) throw std::runtime_error(err);
  hip_check(he, "ray readback");
  return n;
  ABI_GUARD_END(return 0)
}

size_t getCurrentEyeSamples(float* samples, size_t count) {
  ABI_GUARD_BEGIN
  EyeState* e = current_eye();
  if (!e || !e->has_render || !samples) return 0;
  const size_t n = (size_t)std::min<uint64_t>(e->rendered_rays, count);
  if (n) hip_check(hipMemcpy(samples, e->d_samples, sizeof(float4) * n, hipMemcpyDeviceToHost), "readback");
  return n;
  ABI_GUARD_END(return 0)
}

int probePacketGuard(const float* rays64, int activeLanes, int32_t* ids64) {
  ABI_GUARD_BEGIN
  State& s = st();
  if (!s.loaded) throw std::runtime_error("probePacketGuard: no scene loaded");
  if (!rays64 || !ids64 || activeLanes < 1 || activeLanes > 64)
    throw std::runtime_error("probePacketGuard: 64 rays, 64 ids and 1..64 active lanes expected");
  Ray* d_rays = nullptr;
  int32_t* d_ids = nullptr;
  std::string err;
  bool ok = false;
  hipError_t he = hipMalloc(&d_rays, sizeof(Ray) * 64);
  if (he == hipSuccess) he = hipMalloc(&d_ids, sizeof(int32_t) * 64);
  if (he == hipSuccess) he = hipMemcpy(d_rays, rays64, sizeof(Ray) * 64, hipMemcpyHostToDevice);
  if (he == hipSuccess) {
    ok = launch_packet_guard_probe(s.scene, d_rays, activeLanes, d_ids, s.stream, &err);
    he = hipStreamSynchronize(s.stream);
  }
  if (ok && he == hipSuccess) he = hipMemcpy(ids64, d_ids, sizeof(int32_t) * 64, hipMemcpyDeviceToHost);
  (void)hipFree(d_rays);
  (void)hipFree(d_ids);
  if (!ok && !err.empty()) throw std::runtime_error(err);
  hip_check(he, "probePacketGuard");
  const bool flagged = __atomic_load_n(s.h_fault, __ATOMIC_ACQUIRE) != 0u;
  *s.h_fault = 0u;
  return flagged ? 1 : 0;
  ABI_GUARD_END(return -1)
}

int traceRays(const float* rays, size_t count, int32_t* primIds, float* tbb) {
  ABI_GUARD_BEGIN
  State& s = st();
  if (!s.loaded) throw std::runtime_error("traceRays: no scene loaded");
  if (count == 0) return 0;
  if (!rays || !primIds || !tbb) throw std::runtime_error("traceRays: null argument");
  Ray* d_rays = nullptr;
  int32_t* d_hit = nullptr;
  float4* d_tbb = nullptr;
  std::string err;
  bool ok = false;
  hipError_t he = hipMalloc(&d_rays, sizeof(Ray) * count);
  if (he == hipSuccess) he = hipMalloc(&d_hit, sizeof(int32_t) * count);
  if (he == hipSuccess) he = hipMalloc(&d_tbb, sizeof(float4) * count);
  if (he == hipSuccess) he = hipMemcpy(d_rays, rays, sizeof(Ray) * count, hipMemcpyHostToDevice);
  if (he == hipSuccess) {
    ok = launch_trace_rays(s.scene, d_rays, count, d_hit, d_tbb, s.stream, &err);
    he = hipStreamSynchronize(s.stream);
  }
  std::vector<float4> t4;
  if (ok && he == hipSuccess) {
    he = hipMemcpy(primIds, d_hit, sizeof(int32_t) * count, hipMemcpyDeviceToHost);
    t4.resize(count);
    if (he == hipSuccess) he = hipMemcpy(t4.data(), d_tbb, sizeof(float4) * count, hipMemcpyDeviceToHost);
  }
  (void)hipFree(d_rays);
  (void)hipFree(d_hit);
  (void)hipFree(d_tbb);
  if (!ok && !err.empty()) throw std::runtime_error(err);
  hip_check(he, "traceRays");
  for (size_t i = 0; i < count; ++i) {
    tbb[3 * i] = t4[i].x;
    tbb[3 * i + 1] = t4[i].y;
    tbb[3 * i + 2] = t4[i].z;
  }
  return 0;
  ABI_GUARD_END(return -1)
}

// Poses of a batch from (x, y, z, rx, ry, rz) rows, posed as setCameraPose does.
static std::vector<Pose> poses_from6(const float* poses, size_t n) {
  std::vector<Pose> ps(n);
  for (size_t k = 0; k < n; ++k) {
    Pose& p = ps[k];
    const float* q = poses + 6 * k;
    pose_reset(&p);
    pose_rotate_around(&p, q[3], 1.0f, 0.0f, 0.0f);
    pose_rotate_around(&p, q[4], 0.0f, 1.0f, 0.0f);
    pose_rotate_around(&p, q[5], 0.0f, 0.0f, 1.0f);
    pose_move(&p, q[0], q[1], q[2]);
  }
  return ps;
}

// Evenly spaced poses of a batch holding about 2^20 units (all of a smaller
// batch): the sample the tuner times.
static std::vector<Pose> tuning_sample(const std::vector<Pose>& ps, size_t omm_count) {
  const size_t want = std::max<size_t>(1, ((size_t)1 << 20) / std::max<size_t>(omm_count, 1));
  if (ps.size() <= want) return ps;
  std::vector<Pose> out(want);
  for (size_t k = 0; k < want; ++k) out[k] = ps[k * ps.size() / want];
  return out;
}

static LaunchShape tune_shape(EyeState& e, UnitLaunch U, const std::vector<Pose>& ps, int repeats,
                              std::vector<TuneTrial>* trials) {
  State& s = st();
  std::string err;
  LaunchShape best;
  if (!s.batch.tune(s.scene, U, tuning_sample(ps, e.omms.size()), repeats, s.stream, trials, &best, &err))
    throw std::runtime_error(err);
  double best_ms = 0.0;
  for (const TuneTrial& t : *trials)
    if (t.shape.units_per_grab == best.units_per_grab && t.shape.blocks == best.blocks &&
        t.shape.traversal == best.traversal)
      best_ms = t.ms;
  s.tuner.put(LaunchTuner::key_of(s.scene, U, e.total_samples()), best, best_ms);
  if (!s.tune_cache.empty() && !s.tuner.save(s.tune_cache, &err)) throw std::runtime_error(err);
  return best;
}

double renderPoseBatch(const float* poses, size_t poseCount, float* rgbOut) {
  ABI_GUARD_BEGIN
  State& s = st();
  EyeState* e = current_eye();
  if (!s.loaded || !e) throw std::runtime_error("renderPoseBatch: no compound eye selected");
  if (poseCount == 0) return 0.0;
  if (!poses) throw std::runtime_error("renderPoseBatch: null poses");
  upload_eye(*e);
  const std::vector<Pose> ps = poses_from6(poses, poseCount);
  UnitLaunch U = unit_launch_of(*e, nullptr, (int)poseCount);
  if (s.autotune) {
    LaunchShape shape;
    if (!s.tuner.find(LaunchTuner::key_of(s.scene, U, e->total_samples()), &shape)) {
      std::vector<TuneTrial> trials;
      shape = tune_shape(*e, U, ps, 1, &trials);
    }
    U.units_per_grab = shape.units_per_grab;
    U.blocks = shape.blocks;
    U.traversal = shape.traversal;
  }
  std::string err;
  double ms = -1.0;
  if (!s.batch.render(s.scene, U, ps, rgbOut, s.stream, &ms, &err)) throw std::runtime_error(err);
  check_packet_fault("renderPoseBatch");
  return ms;
  ABI_GUARD_END(return -1.0)
}

int renderPoseBatchInto(const float* devPoses, size_t poseCount, float* devColours, void* stream) {
  ABI_GUARD_BEGIN
  State& s = st();
  EyeState* e = current_eye();
  if (!s.loaded || !e) throw std::runtime_error("renderPoseBatchInto: no compound eye selected");
  if (poseCount == 0) return 0;
  if (!devPoses || !devColours) throw std::runtime_error("renderPoseBatchInto: null device pointer");
  if ((reinterpret_cast<uintptr_t>(devPoses) & 3u) != 0 || (reinterpret_cast<uintptr_t>(devColours) & 15u) != 0)
    throw std::runtime_error("renderPoseBatchInto: poses must be 4-byte and colours 16-byte aligned");
  const uint64_t units = (uint64_t)poseCount * (uint64_t)e->omms.size();
  if (poseCount > (size_t)INT32_MAX || units > kNoAllocMaxUnits)
    throw std::runtime_error("renderPoseBatchInto: " + std::to_string(units) + " units exceed the maximum of " +
                             std::to_string(kNoAllocMaxUnits) + " per call");
  if (e->device_dirty())
    throw std::runtime_error(
        "renderPoseBatchInto: the current eye's ommatidia or sample counts changed since they were last uploaded; "
        "render the eye once with renderFrame or renderPoseBatch first (those calls may allocate)");
  uint32_t* scratch = no_alloc_scratch(0);
  if (!scratch) throw std::runtime_error("renderPoseBatchInto: fixed scratch not initialised");
  UnitLaunch U = unit_launch_of(*e, reinterpret_cast<const Pose*>(devPoses), (int)poseCount);
  U.colours = reinterpret_cast<float4*>(devColours);
  U.queue = scratch;
  std::string err;
  if (!launch_render_units(s.scene, U, stream ? static_cast<hipStream_t>(stream) : s.stream, &err))
    throw std::runtime_error(err);
  return 0;
  ABI_GUARD_END(return -1)
}

void setBatchAutotune(int enabled) { st().autotune = enabled != 0; }

int64_t getAdaptiveStages(float* pilot, float* topup, uint32_t* extra, size_t maxUnits) {
  ABI_GUARD_BEGIN
  std::string err;
  if (!st().adaptive.copy_last(maxUnits, reinterpret_cast<float4*>(pilot), reinterpret_cast<float4*>(topup), extra,
                               &err))
    throw std::runtime_error(err);
  return (int64_t)maxUnits;
  ABI_GUARD_END(return -1)
}

int setAdaptiveLabelError(float labelError) {
  if (!(labelError > 0.0f)) {
    set_error("setAdaptiveLabelError: the error must be positive");
    return -1;
  }
  st().label_error = labelError;
  return 0;
}

void setPacketCounting(bool enabled) { st().count_packets = enabled; }

int getPacketCounts(uint64_t* out2) {
  ABI_GUARD_BEGIN
  State& s = st();
  if (!out2) throw std::runtime_error("getPacketCounts: null output");
  ensure_hip();
  unsigned long long h[2] = {0, 0};
  hip_check(hipStreamSynchronize(s.stream), "sync");
  hip_check(hipMemcpy(h, s.d_packet_counts, sizeof(h), hipMemcpyDeviceToHost), "packet counts readback");
  hip_check(hipMemset(s.d_packet_counts, 0, sizeof(h)), "clear packet counts");
  out2[0] = h[0];
  out2[1] = h[1];
  return 0;
  ABI_GUARD_END(return -1)
}

int getAdaptiveTimings(double* out4) {
  if (!out4) return -1;
  for (int k = 0; k < 4; ++k) out4[k] = st().adaptive_ms[k];
  return 0;
}

int setAdaptiveSampler(int mode) {
  if (mode != kSamplerNetwork && mode != kSamplerPooledSigma) {
    set_error("setAdaptiveSampler: mode must be 0 (network) or 1 (pooled pilot deviation)");
    return -1;
  }
  st().sampler = mode;
  return 0;
}

int setAutotuneCacheFile(const char* path) {
  ABI_GUARD_BEGIN
  State& s = st();
  s.tune_cache = path ? path : "";
  std::string err;
  // merge the file's entries with the tuner's, and keep the file up to date
  if (!s.tune_cache.empty() && (!s.tuner.load(s.tune_cache, &err) || !s.tuner.save(s.tune_cache, &err)))
    throw std::runtime_error(err);
  return (int)s.tuner.size();
  ABI_GUARD_END(return -1)
}

int autotuneRenderBatch(const float* poses, size_t poseCount, int repeats, float* trials, size_t maxTrials) {
  ABI_GUARD_BEGIN
  State& s = st();
  EyeState* e = current_eye();
  if (!s.loaded || !e) throw std::runtime_error("autotuneRenderBatch: no compound eye selected");
  if (!poses || poseCount == 0) throw std::runtime_error("autotuneRenderBatch: no poses");
  upload_eye(*e);
  const std::vector<Pose> ps = poses_from6(poses, poseCount);
  UnitLaunch U = unit_launch_of(*e, nullptr, (int)poseCount);
  std::vector<TuneTrial> tr;
  const LaunchShape best = tune_shape(*e, U, ps, std::max(repeats, 1), &tr);
  check_packet_fault("autotuneRenderBatch");
  // the winner becomes the shape of every later unit launch
  s.units_per_grab = best.units_per_grab;
  s.unit_blocks = best.blocks;
  set_trace_mode(trace_mode_of(best.traversal));
  if (trials)
    for (size_t i = 0; i < tr.size() && i < maxTrials; ++i) {
      trials[4 * i + 0] = (float)tr[i].shape.units_per_grab;
      trials[4 * i + 1] = (float)tr[i].shape.blocks;
      trials[4 * i + 2] = (float)tr[i].shape.traversal;
      trials[4 * i + 3] = (float)tr[i].ms;
    }
  return (int)tr.size();
  ABI_GUARD_END(return -1)
}

int getRenderTuning(int* shape) {
  if (!shape) return -1;
  State& s = st();
  shape[0] = s.units_per_grab;
  shape[1] = s.unit_blocks;
  shape[2] = trace_mode();
  return 0;
}

size_t mlpParamCount(void) { return (size_t)kMlpParams; }

int mlpInit(uint32_t seed) {
  ABI_GUARD_BEGIN
  ensure_hip();
  std::string err;
  if (!st().mlp.init(seed, st().stream, &err)) throw std::runtime_error(err);
  return 0;
  ABI_GUARD_END(return -1)
}

int mlpSetParams(const float* params, size_t count) {
  ABI_GUARD_BEGIN
  if (!params || count != (size_t)kMlpParams) throw std::runtime_error("mlpSetParams: need mlpParamCount() floats");
  ensure_hip();
  std::string err;
  if (!st().mlp.set_params(params, st().stream, &err)) throw std::runtime_error(err);
  return 0;
  ABI_GUARD_END(return -1)
}

int mlpGetParams(float* params, size_t count) {
  ABI_GUARD_BEGIN
  if (!params || count != (size_t)kMlpParams) throw std::runtime_error("mlpGetParams: need mlpParamCount() floats");
  ensure_hip();
  std::string err;
  if (!st().mlp.get_params(params, st().stream, &err)) throw std::runtime_error(err);
  return 0;
  ABI_GUARD_END(return -1)
}

int mlpPredict(const float* features, size_t rows, float* out) {
  ABI_GUARD_BEGIN
  if (rows == 0) return 0;
  if (!features || !out) throw std::runtime_error("mlpPredict: null argument");
  ensure_hip();
  DeviceTemp tmp;
  const float* dx = tmp.upload(features, rows * kMlpIn);
  float* dy = tmp.upload<float>(nullptr, rows);
  std::string err;
  if (!st().mlp.predict(dx, (int64_t)rows, dy, st().stream, &err)) throw std::runtime_error(err);
  hip_check(hipStreamSynchronize(st().stream), "mlp");
  hip_check(hipMemcpy(out, dy, sizeof(float) * rows, hipMemcpyDeviceToHost), "readback");
  return 0;
  ABI_GUARD_END(return -1)
}

int mlpLossGrad(const float* features, const float* targets, size_t rows, float* loss, float* grad) {
  ABI_GUARD_BEGIN
  if (!features || !targets || rows == 0) throw std::runtime_error("mlpLossGrad: empty input");
  ensure_hip();
  DeviceTemp tmp;
  const float* dx = tmp.upload(features, rows * kMlpIn);
  const float* dt = tmp.upload(targets, rows);
  float* dl = tmp.upload<float>(nullptr, 1);
  float* dg = tmp.upload<float>(nullptr, kMlpParams);
  std::string err;
  if (!st().mlp.loss_grad(dx, dt, (int64_t)rows, dl, dg, st().stream, &err)) throw std::runtime_error(err);
  hip_check(hipStreamSynchronize(st().stream), "mlp");
  if (loss) hip_check(hipMemcpy(loss, dl, sizeof(float), hipMemcpyDeviceToHost), "readback");
  if (grad) hip_check(hipMemcpy(grad, dg, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost), "readback");
  return 0;
  ABI_GUARD_END(return -1)
}

int mlpTrain(const float* features, const float* targets, size_t rows, int epochs, int batch, float lr,
             uint32_t seed, float* epochLosses) {
  ABI_GUARD_BEGIN
  if (!features || !targets || rows == 0) throw std::runtime_error("mlpTrain: empty input");
  ensure_hip();
  DeviceTemp tmp;
  const float* dx = tmp.upload(features, rows * kMlpIn);
  const float* dt = tmp.upload(targets, rows);
  std::vector<float> losses;
  std::string err;
  if (!st().mlp.train(dx, dt, (int64_t)rows, epochs, batch, lr, seed, &losses, st().stream, &err))
    throw std::runtime_error(err);
  if (epochLosses) std::copy(losses.begin(), losses.end(), epochLosses);
  return 0;
  ABI_GUARD_END(return -1)
}


int nearestOmmatidia(const float* ommatidia, size_t count, int k, uint32_t* out) {
  ABI_GUARD_BEGIN
  if (k < 1 || k > 16) throw std::runtime_error("nearestOmmatidia: k must be 1..16");
  if (count && (!ommatidia || !out)) throw std::runtime_error("nearestOmmatidia: null array");
  nearest_ommatidia(ommatidia, count, k, kNeighbourPosScale, out);
  return 0;
  ABI_GUARD_END(return -1)
}

int64_t adaptiveTrainingSet(const float* poses, size_t poseCount, const uint8_t* unitMask, int pilot, int labelSpp,
                            float errorTarget, float* features, float* labels, size_t maxRows) {
  ABI_GUARD_BEGIN
  State& s = st();
  EyeState* e = current_eye();
  if (!s.loaded || !e) throw std::runtime_error("adaptiveTrainingSet: no compound eye selected");
  if (!poses && poseCount) throw std::runtime_error("adaptiveTrainingSet: null poses");
  upload_eye(*e);
  const size_t units = poseCount * e->omms.size();
  std::vector<uint8_t> mask(units, 1);
  if (unitMask) mask.assign(unitMask, unitMask + units);
  std::vector<float> f, l;
  std::string err;
  UnitLaunch U = unit_launch_of(*e, nullptr, (int)poseCount);
  if (!s.adaptive.training_set(s.scene, U, poses_from6(poses, poseCount), mask, pilot, labelSpp, errorTarget,
                               neighbours_of(*e), &f, &l, s.stream, &err))
    throw std::runtime_error(err);
  check_packet_fault("adaptiveTrainingSet");
  s.label_error = errorTarget;
  const size_t rows = std::min(l.size(), maxRows);
  if (features) std::copy(f.begin(), f.begin() + (ptrdiff_t)(rows * kMlpIn), features);
  if (labels) std::copy(l.begin(), l.begin() + (ptrdiff_t)rows, labels);
  return (int64_t)rows;
  ABI_GUARD_END(return -1)
}

double renderPoseBatchAdaptive(const float* poses, size_t poseCount, float errorTarget, int maxSpp,
                               float budgetScale, float* rgbOut, uint32_t* countsOut, uint64_t* totalSamples) {
  ABI_GUARD_BEGIN
  State& s = st();
  EyeState* e = current_eye();
  if (!s.loaded || !e) throw std::runtime_error("renderPoseBatchAdaptive: no compound eye selected");
  if (!poses && poseCount) throw std::runtime_error("renderPoseBatchAdaptive: null poses");
  upload_eye(*e);
  AdaptiveConfig cfg;
  cfg.error_target = errorTarget;
  cfg.label_error = s.label_error;
  cfg.max_spp = maxSpp;
  cfg.budget_scale = budgetScale;
  cfg.neighbours = neighbours_of(*e);
  cfg.sampler = s.sampler;
  AdaptiveResult res;
  std::string err;
  UnitLaunch U = unit_launch_of(*e, nullptr, (int)poseCount);
  if (!s.adaptive.render(s.scene, U, poses_from6(poses, poseCount), s.mlp, cfg, rgbOut, countsOut, &res, s.stream,
                         &err))
    throw std::runtime_error(err);
  check_packet_fault("renderPoseBatchAdaptive");
  for (int k = 0; k < 4; ++k) s.adaptive_ms[k] = res.stage_ms[k];
  if (totalSamples) *totalSamples = res.total_samples;
  return res.gpu_ms;
  ABI_GUARD_END(return -1.0)
}

int getSceneBVHInfo(int64_t* info) {
  State& s = st();
  if (!info || !s.loaded) return -1;
  info[0] = s.scene.tri_count;
  info[1] = s.scene.node_count;
  info[2] = s.scene.stack_bound;
  info[3] = s.scene.levels;
  info[4] = s.scene.leaf_max;
  info[5] = (int64_t)(s.scene.build_ms * 1000.0);
  info[6] = (int64_t)sizeof(WideNode);
  info[7] = 0;
  return 0;
}

size_t copySceneBVHNodes(void* dst, size_t bytes) {
  ABI_GUARD_BEGIN
  State& s = st();
  if (!s.loaded || !dst) return 0;
  const size_t n = std::min(bytes / sizeof(WideNode), (size_t)s.scene.node_count);
  if (n) hip_check(hipMemcpy(dst, s.scene.nodes, n * sizeof(WideNode), hipMemcpyDeviceToHost), "readback");
  return n * sizeof(WideNode);
  ABI_GUARD_END(return 0)
}

size_t copySceneSortOrder(uint32_t* ids, uint32_t* codes, size_t count) {
  ABI_GUARD_BEGIN
  State& s = st();
  if (!s.loaded) return 0;
  const size_t n = std::min(count, (size_t)s.scene.tri_count);
  if (n && ids) hip_check(hipMemcpy(ids, s.scene.sorted_ids, n * 4, hipMemcpyDeviceToHost), "readback");
  if (n && codes) hip_check(hipMemcpy(codes, s.scene.codes, n * 4, hipMemcpyDeviceToHost), "readback");
  return n;
  ABI_GUARD_END(return 0)
}

int setSceneLeafSize(int leafSize) {
  ABI_GUARD_BEGIN
  State& s = st();
  if (leafSize < 1 || leafSize > kLeafMaxLimit) throw std::runtime_error("setSceneLeafSize: leaf size must be 1..16");
  s.leaf_max = leafSize;
  if (s.loaded) build_scene_bvh();
  return 0;
  ABI_GUARD_END(return -1)
}

void setTraversalMode(int mode) {
  set_trace_mode(mode == 0 ? kTracePerRay : mode == 2 ? kTracePacketScalar : kTracePacket);
}

void setSampleRecording(bool enabled) { st().record_samples = enabled; }

const char* getLastRenderKernel(void) { return last_render_kernel(); }

void setSplitLaunches(bool enabled) { set_split_launches(enabled); }

void setTaskSplit(int mode) {
  if (mode < 0 || mode > 2) {
    set_error("setTaskSplit: mode must be 0 (never), 1 (always) or 2 (by measurement)");
    return;
  }
  if (mode != g_task_split) {
    g_task_split = mode;
    for (auto& e : st().eyes) split_restart(e);
  }
}

int getCurrentEyeTaskSplits(void) {
  EyeState* e = current_eye();
  return e ? (int)(e->split_listed[0] + e->split_listed[1]) : 0;
}

void setRenderTuning(int unitsPerGrab, int blocks) {
  State& s = st();
  s.units_per_grab = std::max(0, std::min(unitsPerGrab, kMaxUnitsPerGrab));
  s.unit_blocks = std::max(0, blocks);
}

double runCopyProbe(size_t bytes, int repeats) {
  ABI_GUARD_BEGIN
  ensure_hip();
  const size_t n = bytes / sizeof(float4);
  if (n == 0 || repeats < 1) return 0.0;
  DeviceTemp tmp;
  float4* src = tmp.upload<float4>(nullptr, n);
  float4* dst = tmp.upload<float4>(nullptr, n);
  State& s = st();
  hip_check(hipMemsetAsync(src, 0, n * sizeof(float4), s.stream), "memset");
  std::string err;
  if (!launch_copy_probe(src, dst, n, s.stream, &err)) throw std::runtime_error(err);  // warm-up
  hip_check(hipEventRecord(s.ev[0], s.stream), "event");
  for (int r = 0; r < repeats; ++r)
    if (!launch_copy_probe(src, dst, n, s.stream, &err)) throw std::runtime_error(err);
  hip_check(hipEventRecord(s.ev[1], s.stream), "event");
  hip_check(hipEventSynchronize(s.ev[1]), "sync");
  float ms = 0.0f;
  hip_check(hipEventElapsedTime(&ms, s.ev[0], s.ev[1]), "elapsed");
  return ms / repeats;
  ABI_GUARD_END(return -1.0)
}

int64_t parseGlTFTriangles(const char* path, float* pos, float* col, uint32_t* mode, size_t maxTris) {
  ABI_GUARD_BEGIN
  if (!path) throw std::runtime_error("parseGlTFTriangles: null path");
  HostScene hs;
  std::string err;
  if (!load_gltf(path, &hs, &err)) throw std::runtime_error(err);
  const size_t n = hs.tri_count();
  const size_t m = std::min(n, maxTris);
  if (m) {
    if (pos) memcpy(pos, hs.pos.data(), sizeof(float) * 9 * m);
    if (col) memcpy(col, hs.col.data(), sizeof(float) * 9 * m);
    if (mode) memcpy(mode, hs.mode.data(), sizeof(uint32_t) * m);
  }
  return (int64_t)n;
  ABI_GUARD_END(return -1)
}

size_t getCurrentCameraRays(float* rays, size_t maxRays) {
  ABI_GUARD_BEGIN
  State& s = st();
  Cam* c = current_cam();
  if (!s.loaded || !c || c->eye >= 0) return 0;
  const size_t n = (size_t)std::max(s.width, 0) * (size_t)std::max(s.height, 0);
  if (!rays || maxRays == 0) return n;
  DeviceTemp tmp;
  Ray* d = tmp.upload<Ray>(nullptr, n);
  std::string err;
  if (!launch_camera_frame(s.scene, camera_launch_of(*c), nullptr, d, s.stream, &err)) throw std::runtime_error(err);
  hip_check(hipStreamSynchronize(s.stream), "camera rays");
  const size_t m = std::min(n, maxRays);
  hip_check(hipMemcpy(rays, d, sizeof(Ray) * m, hipMemcpyDeviceToHost), "readback");
  return m;
  ABI_GUARD_END(return 0)
}

int64_t parseGlTFTextures(const char* path, int32_t* desc, size_t maxTextures, uint32_t* texels, size_t maxTexels,
                          int64_t* texelCount) {
  ABI_GUARD_BEGIN
  if (!path) throw std::runtime_error("parseGlTFTextures: null path");
  HostScene hs;
  std::string err;
  if (!load_gltf(path, &hs, &err)) throw std::runtime_error(err);
  const size_t nt = std::min(hs.textures.size(), maxTextures);
  if (desc)
    for (size_t t = 0; t < nt; ++t) {
      const TextureDesc& d = hs.textures[t];
      int32_t* o = desc + 6 * t;
      o[0] = d.width;
      o[1] = d.height;
      o[2] = d.wrap_s;
      o[3] = d.wrap_t;
      o[4] = d.filter;
      o[5] = (int32_t)d.first;
    }
  if (texels) memcpy(texels, hs.texels.data(), sizeof(uint32_t) * std::min(hs.texels.size(), maxTexels));
  if (texelCount) *texelCount = (int64_t)hs.texels.size();
  return (int64_t)hs.textures.size();
  ABI_GUARD_END(return -1)
}

}  // extern "C"
