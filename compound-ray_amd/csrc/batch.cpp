#include "batch.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

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
         a.sample_offsets == b.sample_offsets && a.queue == b.queue && a.units_per_grab == b.units_per_grab &&
         a.auto_grab_max == b.auto_grab_max && a.blocks == b.blocks && a.traversal == b.traversal &&
         a.packet_counts == b.packet_counts;
}
}  // namespace

std::vector<LaunchShape> LaunchTuner::candidates(int capacity) {
  std::vector<LaunchShape> out;
  const int cap = std::max(capacity, 1);
  const int block_opts[3] = {std::max(cap / 2, 1), cap, 2 * cap};
  for (int trav = kUnitPacket; trav <= kUnitPerRay; ++trav)
    for (int b : block_opts)
      for (int g : {1, 2, 4, 8, 16}) {
        LaunchShape sh;
        sh.units_per_grab = g;
        sh.blocks = b;
        sh.traversal = trav;
        out.push_back(sh);
      }
  return out;
}

std::string LaunchTuner::key_of(const DeviceScene& scene, const UnitLaunch& U, uint64_t samples_per_pose) {
  char buf[160];
  std::snprintf(buf, sizeof(buf), "tris%d_nodes%d_leaf%d_omm%d_spp%llu_stats%d_rec%d", scene.tri_count,
                scene.node_count, scene.leaf_max, U.omm_count, (unsigned long long)samples_per_pose,
                U.stats ? 1 : 0, U.samples ? 1 : 0);
  return buf;
}

bool LaunchTuner::find(const std::string& key, LaunchShape* shape) const {
  for (size_t i = keys_.size(); i-- > 0;)
    if (keys_[i] == key) {
      *shape = shapes_[i];
      return true;
    }
  return false;
}

void LaunchTuner::put(const std::string& key, const LaunchShape& shape, double ms) {
  for (size_t i = 0; i < keys_.size(); ++i)
    if (keys_[i] == key) {
      shapes_[i] = shape;
      ms_[i] = ms;
      return;
    }
  keys_.push_back(key);
  shapes_.push_back(shape);
  ms_.push_back(ms);
}

bool LaunchTuner::load(const std::string& path, std::string* err) {
  std::ifstream f(path);
  if (!f) return true;  // no file yet: nothing cached
  std::string line;
  int lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    if (line.empty() || line[0] == '#') continue;
    std::istringstream in(line);
    std::string key;
    LaunchShape sh;
    double ms = 0.0;
    if (!(in >> key >> sh.units_per_grab >> sh.blocks >> sh.traversal >> ms) || sh.units_per_grab < 1 ||
        sh.blocks < 0 || sh.traversal < kUnitPacket || sh.traversal > kUnitPerRay) {
      if (err) *err = "autotune cache " + path + ": bad entry on line " + std::to_string(lineno);
      return false;
    }
    put(key, sh, ms);
  }
  return true;
}

bool LaunchTuner::save(const std::string& path, std::string* err) const {
  std::ofstream f(path, std::ios::trunc);
  if (!f) {
    if (err) *err = "autotune cache " + path + ": cannot write";
    return false;
  }
  f << "# key units_per_grab blocks traversal ms\n";
  for (size_t i = 0; i < keys_.size(); ++i)
    f << keys_[i] << ' ' << shapes_[i].units_per_grab << ' ' << shapes_[i].blocks << ' ' << shapes_[i].traversal
      << ' ' << ms_[i] << '\n';
  return bool(f);
}

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
  g_valid_ = false;
  g_scene_ = 0;
}

bool BatchRenderer::build_graph(const DeviceScene& scene, const UnitLaunch& U, hipStream_t stream,
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
  const bool ok = launch_render_units(scene, U, stream, &lerr);
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
  g_valid_ = true;
  g_scene_ = scene.generation;
  return true;
}

bool BatchRenderer::prepare(const UnitLaunch& U, const std::vector<Pose>& poses, hipStream_t stream, bool* realloc,
                            std::string* err) {
  const size_t np = poses.size();
  if (!ev0_ && (hipEventCreate(&ev0_) != hipSuccess || hipEventCreate(&ev1_) != hipSuccess)) {
    if (err) *err = "batch: event creation failed";
    return false;
  }
  const size_t units = np * (size_t)U.omm_count;
  *realloc = poses_cap_ < np || colours_cap_ < units;
  if (!grow(&d_poses_, &poses_cap_, np, err)) return false;
  if (!grow(&d_colours_, &colours_cap_, units, err)) return false;
  if (!d_queue_ && hipMalloc(&d_queue_, sizeof(uint32_t) * kQueueWords) != hipSuccess) {
    if (err) *err = "batch: queue allocation failed";
    return false;
  }
  if (hipMemcpyAsync(d_poses_, poses.data(), sizeof(Pose) * np, hipMemcpyHostToDevice, stream) != hipSuccess) {
    if (err) *err = "batch: pose upload failed";
    return false;
  }
  return true;
}

bool BatchRenderer::tune(const DeviceScene& scene, UnitLaunch U, const std::vector<Pose>& poses, int repeats,
                         hipStream_t stream, std::vector<TuneTrial>* trials, LaunchShape* best, std::string* err) {
  trials->clear();
  if (poses.empty() || U.omm_count == 0) {
    if (err) *err = "autotune: empty batch";
    return false;
  }
  bool realloc = false;
  if (!prepare(U, poses, stream, &realloc, err)) return false;
  U.poses = d_poses_;
  U.pose_count = (int)poses.size();
  U.pose_base = 0;
  U.colours = d_colours_;
  U.queue = d_queue_;
  U.stats = nullptr;
  U.samples = nullptr;
  const int capacity = render_units_capacity(scene, U);
  double best_ms = 1e300;
  for (const LaunchShape& sh : LaunchTuner::candidates(capacity)) {
    U.units_per_grab = sh.units_per_grab;
    U.blocks = sh.blocks;
    U.traversal = sh.traversal;
    if (!launch_render_units(scene, U, stream, err)) return false;  // untimed: caches, clocks
    double ms_min = 1e300;
    for (int r = 0; r < std::max(repeats, 1); ++r) {
      (void)hipEventRecord(ev0_, stream);
      if (!launch_render_units(scene, U, stream, err)) return false;
      (void)hipEventRecord(ev1_, stream);
      if (hipEventSynchronize(ev1_) != hipSuccess) {
        if (err) *err = "autotune: render failed";
        return false;
      }
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, ev0_, ev1_);
      ms_min = std::min(ms_min, (double)ms);
    }
    trials->push_back(TuneTrial{sh, ms_min});
    if (ms_min < best_ms) {
      best_ms = ms_min;
      *best = sh;
    }
  }
  return true;
}

bool BatchRenderer::render(const DeviceScene& scene, UnitLaunch U, const std::vector<Pose>& poses, float* rgb,
                           hipStream_t stream, double* gpu_ms, std::string* err) {
  const size_t np = poses.size();
  *gpu_ms = 0.0;
  if (np == 0) return true;
  const size_t units = np * (size_t)U.omm_count;
  bool realloc = false;
  if (!prepare(U, poses, stream, &realloc, err)) return false;
  U.poses = d_poses_;
  U.pose_count = (int)np;
  U.pose_base = 0;
  U.colours = d_colours_;
  U.queue = d_queue_;
  if (use_graph_) {
    if (realloc || !exec_ || !g_valid_ || g_scene_ != scene.generation || !same_launch(U, g_launch_))
      if (!build_graph(scene, U, stream, err)) return false;
    (void)hipEventRecord(ev0_, stream);
    if (hipGraphLaunch(exec_, stream) != hipSuccess) {
      if (err) *err = "batch: graph launch failed";
      return false;
    }
  } else {
    (void)hipEventRecord(ev0_, stream);
    if (!launch_render_units(scene, U, stream, err)) return false;
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
