// Learned adaptive sampling on gfx950: pilot statistics -> features -> MLP ->
// per-pose Neyman allocation -> top-up render -> combination.  Every step is a
// kernel on one stream (no host round trip inside a batch).
#include "adaptive.h"

#include <algorithm>
#include <cmath>

namespace eye {
namespace {

constexpr int kThreads = 256;

__global__ void k_features(const Ommatidium* __restrict__ omms, const Pose* __restrict__ poses, int omm_count,
                           uint64_t units, const float4* __restrict__ col, const float4* __restrict__ stats,
                           const uint32_t* __restrict__ nbr, float* __restrict__ feat) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= units) return;
  const int p = (int)(u / (uint64_t)omm_count);
  const int i = (int)(u - (uint64_t)p * (uint64_t)omm_count);
  const Ommatidium o = omms[i];
  const Pose P = poses[p];
  const V3 a = normalize(v3(o.dx, o.dy, o.dz));
  const V3 w = normalize((v3(P.x[0], P.x[1], P.x[2]) * a.x + v3(P.y[0], P.y[1], P.y[2]) * a.y) +
                         v3(P.z[0], P.z[1], P.z[2]) * a.z);
  const float4 c = col[u];
  const float4 s0 = stats[2 * u];
  const float4 s1 = stats[2 * u + 1];
  float* f = feat + u * kMlpIn;
  f[0] = w.x;
  f[1] = w.y;
  f[2] = w.z;
  f[3] = o.acceptance * 20.0f;
  f[4] = c.x;
  f[5] = c.y;
  f[6] = c.z;
  f[7] = sqrtf(s0.x);
  f[8] = sqrtf(s0.y);
  f[9] = sqrtf(s0.z);
  f[10] = s1.x;
  f[11] = sqrtf(s1.y);
  // pooled pilot statistics of the nearest ommatidia at the same pose
  const uint64_t row = (uint64_t)p * (uint64_t)omm_count;
  float s_sum = 0.0f, s_max = 0.0f, contrast = 0.0f, d_sum = 0.0f;
  for (int m = 0; m < kNeighbours; ++m) {
    const uint64_t v = row + nbr[(size_t)i * kNeighbours + m];
    const float4 cv = col[v];
    const float4 a = stats[2 * v];
    const float4 b = stats[2 * v + 1];
    const float sv = (sqrtf(a.x) + sqrtf(a.y) + sqrtf(a.z)) * (1.0f / 3.0f);
    s_sum += sv;
    s_max = fmaxf(s_max, sv);
    contrast += (fabsf(cv.x - c.x) + fabsf(cv.y - c.y) + fabsf(cv.z - c.z)) * (1.0f / 3.0f);
    d_sum += sqrtf(b.y);
  }
  constexpr float kInvK = 1.0f / (float)kNeighbours;
  f[12] = s_sum * kInvK;
  f[13] = s_max;
  f[14] = contrast * kInvK;
  f[15] = d_sum * kInvK;
}

__device__ inline uint32_t lane_footprint(uint32_t e) {
  if (e == 0u) return 0u;
  if (e > 32u) return (e + 63u) & ~63u;
  uint32_t p = 1u;
  while (p < e) p <<= 1;
  return p;
}

__device__ inline float required_samples(float log2_nreq) {
  return exp2f(fminf(fmaxf(log2_nreq, 0.0f), 20.0f));
}

// Allocation in three steps, every step spread over the whole GPU (a batch
// can hold one pose of millions of units) and summed in a fixed order, so the
// counts do not depend on scheduling:
//   1. k_alloc_partial: block (pose, c) sums sqrt(n_req) over units
//      [c * kAllocChunk, (c + 1) * kAllocChunk) of the pose;
//   2. k_alloc_total: S[pose] = the sum of the pose's partial sums;
//   3. k_allocate: every unit's count from S[pose], and the total.
constexpr int kAllocChunk = 4096;

__device__ inline float block_sum(float v, float* part) {
  part[threadIdx.x] = v;
  __syncthreads();
  for (int w = kThreads / 2; w >= 1; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  return part[0];
}

__global__ __launch_bounds__(kThreads) void k_alloc_partial(const float* __restrict__ pred, int omm_count, int chunks,
                                                           float* __restrict__ partial) {
  __shared__ float part[kThreads];
  const int pose = (int)(blockIdx.x / (unsigned)chunks);
  const int c = (int)(blockIdx.x - (unsigned)pose * (unsigned)chunks);
  const uint64_t base = (uint64_t)pose * (uint64_t)omm_count;
  const int i1 = min(omm_count, (c + 1) * kAllocChunk);
  float s = 0.0f;
  for (int i = c * kAllocChunk + (int)threadIdx.x; i < i1; i += kThreads) s += sqrtf(required_samples(pred[base + i]));
  const float t = block_sum(s, part);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(kThreads) void k_alloc_total(const float* __restrict__ partial, int chunks,
                                                         float* __restrict__ total_sqrt) {
  __shared__ float part[kThreads];
  float s = 0.0f;
  for (int c = (int)threadIdx.x; c < chunks; c += kThreads) s += partial[(size_t)blockIdx.x * chunks + c];
  const float t = block_sum(s, part);
  if (threadIdx.x == 0) total_sqrt[blockIdx.x] = t;
}

// n_i = sqrt(n_req,i) * S / N * scale, clamped to [pilot, max_spp].
__global__ __launch_bounds__(kThreads) void k_allocate(const float* __restrict__ pred, int omm_count, uint64_t units,
                                                      const float* __restrict__ total_sqrt, int pilot, int max_spp,
                                                      float scale, uint32_t* __restrict__ counts,
                                                      uint32_t* __restrict__ extra,
                                                      unsigned long long* __restrict__ total) {
  __shared__ unsigned long long part[kThreads];
  const uint64_t u = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  unsigned long long mine = 0;
  if (u < units) {
    const uint64_t pose = u / (uint64_t)omm_count;
    const float S = total_sqrt[pose];
    const float want = sqrtf(required_samples(pred[u])) * S / (float)omm_count * scale;
    const float c = fminf(fmaxf(ceilf(want), (float)pilot), (float)max_spp);
    // The top-up pass gives a unit of e samples a power-of-two lane segment
    // (e <= 32) or whole 64-lane rounds (e > 32); those lanes cost the same
    // time idle or busy, so the top-up is rounded up to fill them.
    const uint32_t e = lane_footprint((uint32_t)c - (uint32_t)pilot);
    const uint32_t n = min((uint32_t)pilot + e, (uint32_t)max_spp);
    counts[u] = n;
    extra[u] = n - (uint32_t)pilot;
    mine = n;
  }
  part[threadIdx.x] = mine;
  __syncthreads();
  for (int w = kThreads / 2; w >= 1; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(total, part[0]);
}

// Final colour of each unit into `out`: the pilot mean alone, or pilot and
// top-up means weighted by their sample counts (the top-up buffer is kept).
__global__ void k_combine(const float4* __restrict__ a, const float4* __restrict__ b,
                          const uint32_t* __restrict__ extra, int pilot, uint64_t units, float4* __restrict__ out) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= units) return;
  const float4 ca = a[u];
  const uint32_t e = extra[u];
  if (e == 0) {
    out[u] = make_float4(ca.x, ca.y, ca.z, (float)pilot);
    return;
  }
  const float4 cb = b[u];
  const float wa = (float)pilot, wb = (float)e, inv = 1.0f / (wa + wb);
  out[u] = make_float4((wa * ca.x + wb * cb.x) * inv, (wa * ca.y + wb * cb.y) * inv, (wa * ca.z + wb * cb.z) * inv,
                       wa + wb);
}
// Training labels from a long render's statistics: log2(max(n_req, 1)), n_req
// = mean RGB variance / error^2, clamped to [0, 12].
__global__ void k_labels(const float4* __restrict__ stats, uint64_t units, float inv_err2, float* __restrict__ lab) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= units) return;
  const float4 s = stats[2 * u];
  const float var = (s.x + s.y + s.z) * (1.0f / 3.0f);
  const float n = fmaxf(var * inv_err2, 1.0f);
  lab[u] = fminf(log2f(n), 12.0f);
}

// Network-free n_req: log2 of the squared pooled pilot standard deviation
// over label_error^2, from the unit's features (own RGB deviations 7-9 and the
// neighbours' mean deviation 12).
__global__ void k_pooled_sigma(const float* __restrict__ feat, uint64_t units, float inv_err2,
                               float* __restrict__ pred) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= units) return;
  const float* f = feat + u * kMlpIn;
  const float own = (f[7] + f[8] + f[9]) * (1.0f / 3.0f);
  const float sigma = (own + (float)kNeighbours * f[12]) * (1.0f / (float)(kNeighbours + 1));
  pred[u] = log2f(fmaxf(sigma * sigma * inv_err2, 1.0f));
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

template <typename T>
bool grow(T** p, size_t n, std::string* err) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  if (n == 0) return true;
  const hipError_t e = hipMalloc(p, sizeof(T) * n);
  if (e != hipSuccess) {
    if (err) *err = std::string("adaptive: hipMalloc: ") + hipGetErrorString(e);
    return false;
  }
  return true;
}

bool ok_launch(const char* what, std::string* err) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    if (err) *err = std::string("adaptive: ") + what + ": " + hipGetErrorString(e);
    return false;
  }
  return true;
}

}  // namespace

void AdaptiveRenderer::release() {
  for (void* p : {(void*)d_poses_, (void*)d_col_a_, (void*)d_col_b_, (void*)d_final_, (void*)d_stats_, (void*)d_feat_,
                  (void*)d_pred_, (void*)d_counts_, (void*)d_extra_, (void*)d_total_, (void*)d_queue_,
                  (void*)d_part_, (void*)d_sqrt_})
    if (p) (void)hipFree(p);
  d_poses_ = nullptr;
  d_part_ = d_sqrt_ = nullptr;
  part_cap_ = 0;
  d_col_a_ = d_col_b_ = d_final_ = d_stats_ = nullptr;
  d_feat_ = d_pred_ = nullptr;
  d_counts_ = d_extra_ = d_queue_ = nullptr;
  d_total_ = nullptr;
  poses_cap_ = units_cap_ = 0;
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  ev0_ = ev1_ = nullptr;
  for (hipEvent_t& e : ev_stage_) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
}

bool AdaptiveRenderer::ensure(size_t poses, size_t units, std::string* err) {
  if (!ev0_ && (hipEventCreate(&ev0_) != hipSuccess || hipEventCreate(&ev1_) != hipSuccess ||
                hipEventCreate(&ev_stage_[0]) != hipSuccess || hipEventCreate(&ev_stage_[1]) != hipSuccess ||
                hipEventCreate(&ev_stage_[2]) != hipSuccess)) {
    if (err) *err = "adaptive: event creation failed";
    return false;
  }
  if (!d_total_ && (!grow(&d_total_, 1, err) || !grow(&d_queue_, kQueueWords, err))) return false;
  if (poses > poses_cap_) {
    if (!grow(&d_poses_, poses, err) || !grow(&d_sqrt_, poses, err)) return false;
    poses_cap_ = poses;
  }
  // per-pose partial sums of the allocation: ceil(units per pose / chunk) each
  const size_t parts = units == 0 ? 0 : poses * ((units / poses + kAllocChunk - 1) / kAllocChunk);
  if (parts > part_cap_) {
    if (!grow(&d_part_, parts, err)) return false;
    part_cap_ = parts;
  }
  if (units > units_cap_) {
    if (!grow(&d_col_a_, units, err) || !grow(&d_col_b_, units, err) || !grow(&d_final_, units, err) ||
        !grow(&d_stats_, 2 * units, err) ||
        !grow(&d_feat_, units * kMlpIn, err) || !grow(&d_pred_, units, err) || !grow(&d_counts_, units, err) ||
        !grow(&d_extra_, units, err))
      return false;
    units_cap_ = units;
  }
  return true;
}

bool AdaptiveRenderer::render(const DeviceScene& scene, UnitLaunch U, const std::vector<Pose>& poses, Mlp& mlp,
                              const AdaptiveConfig& cfg, float* rgb, uint32_t* counts, AdaptiveResult* res,
                              hipStream_t stream, std::string* err) {
  const size_t np = poses.size();
  const size_t units = np * (size_t)U.omm_count;
  if (units == 0) return true;
  if (cfg.sampler != kSamplerNetwork && cfg.sampler != kSamplerPooledSigma) {
    if (err) *err = "adaptive: unknown sampler";
    return false;
  }
  if (cfg.sampler == kSamplerNetwork && !mlp.ready()) {
    if (err) *err = "adaptive: the sampler network is not initialised (mlpInit / mlpTrain)";
    return false;
  }
  if (cfg.pilot < 1 || cfg.max_spp < cfg.pilot || !(cfg.error_target > 0.0f) || !(cfg.label_error > 0.0f)) {
    if (err) *err = "adaptive: need 1 <= pilot <= max samples and positive error targets";
    return false;
  }
  if (!cfg.neighbours) {
    if (err) *err = "adaptive: no neighbour table";
    return false;
  }
  if (!ensure(np, units, err)) return false;
  if (hipMemcpyAsync(d_poses_, poses.data(), sizeof(Pose) * np, hipMemcpyHostToDevice, stream) != hipSuccess) {
    if (err) *err = "adaptive: pose upload failed";
    return false;
  }
  (void)hipEventRecord(ev0_, stream);
  UnitLaunch A = U;
  A.poses = d_poses_;
  A.pose_count = (int)np;
  A.pose_base = 0;
  A.spp = cfg.pilot;
  A.omm_counts = nullptr;
  A.unit_counts = nullptr;
  A.sample_base = 0;
  A.colours = d_col_a_;
  A.stats = d_stats_;
  A.samples = nullptr;
  A.queue = d_queue_;
  // Grab size chosen per pass (0: automatic).  A tuned size comes from uniform
  // batches; a grab of few pilot units would leave most of the wave's lanes idle.
  // Pilot grabs of at most 8 units (one 64-lane round of 8-sample units):
  // on the 256-eye batch 9.6 -> 8.9 ms against 16, the bee eye unchanged.
  A.units_per_grab = 0;
  A.auto_grab_max = 8;
  if (!launch_render_units(scene, A, stream, err)) return false;
  (void)hipEventRecord(ev_stage_[0], stream);
  const unsigned gb = (unsigned)((units + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_features, dim3(gb), dim3(kThreads), 0, stream, U.omms, d_poses_, U.omm_count, (uint64_t)units,
                     d_col_a_, d_stats_, cfg.neighbours, d_feat_);
  if (!ok_launch("features", err)) return false;
  if (cfg.sampler == kSamplerNetwork) {
    if (!mlp.predict(d_feat_, (int64_t)units, d_pred_, stream, err)) return false;
  } else {
    hipLaunchKernelGGL(k_pooled_sigma, dim3(gb), dim3(kThreads), 0, stream, d_feat_, (uint64_t)units,
                       1.0f / (cfg.label_error * cfg.label_error), d_pred_);
    if (!ok_launch("pooled sigma", err)) return false;
  }
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, stream, d_total_);
  const int chunks = (U.omm_count + kAllocChunk - 1) / kAllocChunk;
  hipLaunchKernelGGL(k_alloc_partial, dim3((unsigned)(np * (size_t)chunks)), dim3(kThreads), 0, stream, d_pred_,
                     U.omm_count, chunks, d_part_);
  hipLaunchKernelGGL(k_alloc_total, dim3((unsigned)np), dim3(kThreads), 0, stream, d_part_, chunks, d_sqrt_);
  hipLaunchKernelGGL(k_allocate, dim3(gb), dim3(kThreads), 0, stream, d_pred_, U.omm_count, (uint64_t)units, d_sqrt_,
                     cfg.pilot, cfg.max_spp,
                     cfg.budget_scale * (cfg.label_error / cfg.error_target) * (cfg.label_error / cfg.error_target),
                     d_counts_, d_extra_, d_total_);
  if (!ok_launch("allocate", err)) return false;
  (void)hipEventRecord(ev_stage_[1], stream);
  UnitLaunch B = A;
  B.unit_counts = d_extra_;
  B.sample_base = (uint32_t)cfg.pilot;
  B.colours = d_col_b_;
  B.stats = nullptr;
  // Top-up grabs of at most 6 units: a wave's rounds then mix the cones of
  // fewer, closer ommatidia.  Against 8: the 256-eye batch's top-up 13.2 ->
  // 12.6 ms, the bee eye unchanged; 4 and 2 are slower on the bee eye, 16
  // slower on both.
  B.auto_grab_max = 6;
  if (!launch_render_units(scene, B, stream, err)) return false;
  (void)hipEventRecord(ev_stage_[2], stream);
  hipLaunchKernelGGL(k_combine, dim3(gb), dim3(kThreads), 0, stream, d_col_a_, d_col_b_, d_extra_, cfg.pilot,
                     (uint64_t)units, d_final_);
  if (!ok_launch("combine", err)) return false;
  (void)hipEventRecord(ev1_, stream);
  unsigned long long total = 0;
  if (hipMemcpyAsync(&total, d_total_, sizeof(total), hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess) {
    if (err) *err = "adaptive: render failed";
    return false;
  }
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, ev0_, ev1_);
  if (res) {
    res->gpu_ms = ms;
    res->total_samples = total;
    const hipEvent_t marks[5] = {ev0_, ev_stage_[0], ev_stage_[1], ev_stage_[2], ev1_};
    for (int k = 0; k < 4; ++k) {
      float f = 0.0f;
      (void)hipEventElapsedTime(&f, marks[k], marks[k + 1]);
      res->stage_ms[k] = f;
    }
  }
  if (rgb) {
    std::vector<float4> tmp(units);
    if (hipMemcpy(tmp.data(), d_final_, sizeof(float4) * units, hipMemcpyDeviceToHost) != hipSuccess) {
      if (err) *err = "adaptive: readback failed";
      return false;
    }
    for (size_t u = 0; u < units; ++u) {
      rgb[3 * u] = tmp[u].x;
      rgb[3 * u + 1] = tmp[u].y;
      rgb[3 * u + 2] = tmp[u].z;
    }
  }
  if (counts && hipMemcpy(counts, d_counts_, sizeof(uint32_t) * units, hipMemcpyDeviceToHost) != hipSuccess) {
    if (err) *err = "adaptive: readback failed";
    return false;
  }
  return true;
}

bool AdaptiveRenderer::training_set(const DeviceScene& scene, UnitLaunch U, const std::vector<Pose>& poses,
                                    const std::vector<uint8_t>& unit_mask, int pilot, int label_spp,
                                    float error_target, const uint32_t* neighbours, std::vector<float>* features, std::vector<float>* labels,
                                    hipStream_t stream, std::string* err) {
  const size_t np = poses.size();
  const size_t units = np * (size_t)U.omm_count;
  if (units == 0) return true;
  if (unit_mask.size() != units || pilot < 1 || label_spp < 2 || !(error_target > 0.0f) || !neighbours) {
    if (err) *err = "adaptive: bad training-set request";
    return false;
  }
  if (!ensure(np, units, err)) return false;
  if (hipMemcpyAsync(d_poses_, poses.data(), sizeof(Pose) * np, hipMemcpyHostToDevice, stream) != hipSuccess) {
    if (err) *err = "adaptive: pose upload failed";
    return false;
  }
  // The pilot covers every unit (the neighbour features read the pilots of
  // unmasked units too); only the label render is restricted to the mask.
  std::vector<uint32_t> lc(units);
  for (size_t u = 0; u < units; ++u) lc[u] = unit_mask[u] ? (uint32_t)label_spp : 0u;
  if (hipMemcpyAsync(d_extra_, lc.data(), sizeof(uint32_t) * units, hipMemcpyHostToDevice, stream) != hipSuccess) {
    if (err) *err = "adaptive: count upload failed";
    return false;
  }
  UnitLaunch A = U;
  A.poses = d_poses_;
  A.pose_count = (int)np;
  A.pose_base = 0;
  A.spp = pilot;
  A.omm_counts = nullptr;
  A.unit_counts = nullptr;
  A.sample_base = 0;
  A.colours = d_col_a_;
  A.stats = d_stats_;
  A.samples = nullptr;
  A.queue = d_queue_;
  A.units_per_grab = 0;  // automatic, as in render()
  A.auto_grab_max = 8;
  if (!launch_render_units(scene, A, stream, err)) return false;
  const unsigned gb = (unsigned)((units + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_features, dim3(gb), dim3(kThreads), 0, stream, U.omms, d_poses_, U.omm_count, (uint64_t)units,
                     d_col_a_, d_stats_, neighbours, d_feat_);
  if (!ok_launch("features", err)) return false;
  // label render: independent samples pilot .. pilot + label_spp - 1
  UnitLaunch B = A;
  B.unit_counts = d_extra_;
  B.sample_base = (uint32_t)pilot;
  B.colours = d_col_b_;
  if (!launch_render_units(scene, B, stream, err)) return false;
  hipLaunchKernelGGL(k_labels, dim3(gb), dim3(kThreads), 0, stream, d_stats_, (uint64_t)units,
                     1.0f / (error_target * error_target), d_pred_);
  if (!ok_launch("labels", err)) return false;
  std::vector<float> f(units * kMlpIn), l(units);
  if (hipMemcpyAsync(f.data(), d_feat_, sizeof(float) * f.size(), hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipMemcpyAsync(l.data(), d_pred_, sizeof(float) * l.size(), hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess) {
    if (err) *err = "adaptive: training-set readback failed";
    return false;
  }
  for (size_t u = 0; u < units; ++u) {
    if (!unit_mask[u]) continue;
    features->insert(features->end(), f.begin() + (ptrdiff_t)(u * kMlpIn), f.begin() + (ptrdiff_t)((u + 1) * kMlpIn));
    labels->push_back(l[u]);
  }
  return true;
}

bool AdaptiveRenderer::copy_last(size_t units, float4* pilot, float4* topup, uint32_t* counts, std::string* err) {
  if (units > units_cap_) {
    if (err) *err = "adaptive: fewer units were rendered";
    return false;
  }
  if ((pilot && hipMemcpy(pilot, d_col_a_, sizeof(float4) * units, hipMemcpyDeviceToHost) != hipSuccess) ||
      (topup && hipMemcpy(topup, d_col_b_, sizeof(float4) * units, hipMemcpyDeviceToHost) != hipSuccess) ||
      (counts && hipMemcpy(counts, d_extra_, sizeof(uint32_t) * units, hipMemcpyDeviceToHost) != hipSuccess)) {
    if (err) *err = "adaptive: readback failed";
    return false;
  }
  return true;
}

}  // namespace eye
