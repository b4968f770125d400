// Spherical projections of a compound eye's view (no reference counterpart
// for the precomputation; the reference searches every ommatidium for every
// pixel on every frame, libEyeRenderer3/shaders.cu:412-640).
//
// The pixel -> ommatidium map depends only on the eye's ommatidia and the
// window size, so it is computed once (one thread per pixel, ommatidia staged
// through LDS in tiles) and every later frame is a gather of the reduced
// ommatidial colours.
#include <hip/hip_runtime.h>

#include "projection.h"

namespace eye {
namespace {

constexpr int kProjThreads = 256;

// ids[p] = ommatidium closest in angle to pixel p's sphere direction.  The
// search starts from ommatidium 0 and takes a later one only on a strictly
// larger cosine; the split variant considers ommatidium i >= 1 only when its
// position's x has the side of the pixel (x > 0 right half, x < 0 left half).
__global__ __launch_bounds__(kProjThreads) void k_spherical_ids(const Ommatidium* __restrict__ omms, int omm_count,
                                                                int width, int height, int variant,
                                                                uint32_t* __restrict__ ids) {
  __shared__ float tile[kProjThreads][4];
  const int pixels = width * height;
  const int p = blockIdx.x * kProjThreads + threadIdx.x;
  const bool live = p < pixels;
  const int x = live ? p % width : 0, y = live ? p / width : 0;
  const bool split = variant == kSphericalSplit;
  const V3 s = sphere_direction(x, y, width, height, split);
  const float len_s = sqrtf((s.x * s.x + s.y * s.y) + s.z * s.z);
  const bool right = (float)x / (float)width > 0.5f;
  const bool left = (float)x / (float)width < 0.5f;
  float best = 0.0f;
  uint32_t best_i = 0;
  for (int base = 0; base < omm_count; base += kProjThreads) {
    __syncthreads();
    const int j = base + (int)threadIdx.x;
    if (j < omm_count) {
      const Ommatidium o = omms[j];
      const bool use_pos = variant == kSphericalPosition;
      tile[threadIdx.x][0] = use_pos ? o.px : o.dx;
      tile[threadIdx.x][1] = use_pos ? o.py : o.dy;
      tile[threadIdx.x][2] = use_pos ? o.pz : o.dz;
      tile[threadIdx.x][3] = o.px;
    }
    __syncthreads();
    const int n = min(kProjThreads, omm_count - base);
    for (int k = 0; k < n; ++k) {
      const int i = base + k;
      const float c = cos_angle(v3(tile[k][0], tile[k][1], tile[k][2]), s, len_s);
      if (i == 0) {
        best = c;
        continue;
      }
      if (split && !((tile[k][3] > 0.0f && right) || (tile[k][3] < 0.0f && left))) continue;
      if (c > best) {
        best = c;
        best_i = (uint32_t)i;
      }
    }
  }
  if (live) ids[p] = best_i;
}

__global__ void k_frame_lookup(const float4* __restrict__ colours, const uint32_t* __restrict__ ids, int pixels,
                               int encode_ids, uchar4* __restrict__ frame) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pixels) return;
  const uint32_t i = ids[p];
  if (encode_ids) {
    frame[p] = make_uchar4((unsigned char)(i >> 24), (unsigned char)((i >> 16) & 0xffu),
                           (unsigned char)((i >> 8) & 0xffu), (unsigned char)(i & 0xffu));
    return;
  }
  const float4 c = colours[i];
  frame[p] = make_uchar4(encode_channel(c.x), encode_channel(c.y), encode_channel(c.z), 255);
}

bool ok(const char* what, std::string* err) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    if (err) *err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
  }
  return true;
}

}  // namespace

bool launch_spherical_ids(const Ommatidium* omms, int omm_count, int width, int height, int variant, uint32_t* ids,
                          hipStream_t stream, std::string* err) {
  if (width <= 0 || height <= 0 || omm_count <= 0) return true;
  if (variant < kSphericalPosition || variant > kSphericalSplit) {
    if (err) *err = "spherical projection: unknown variant";
    return false;
  }
  const long long pixels = (long long)width * height;
  if (pixels > 0x7fffffffLL) {
    if (err) *err = "spherical projection: window too large";
    return false;
  }
  hipLaunchKernelGGL(k_spherical_ids, dim3((unsigned)((pixels + kProjThreads - 1) / kProjThreads)),
                     dim3(kProjThreads), 0, stream, omms, omm_count, width, height, variant, ids);
  return ok("spherical_ids", err);
}

bool launch_frame_lookup(const float4* colours, const uint32_t* ids, int pixels, bool encode_ids, uchar4* frame,
                         hipStream_t stream, std::string* err) {
  if (pixels <= 0) return true;
  hipLaunchKernelGGL(k_frame_lookup, dim3((unsigned)((pixels + 255) / 256)), dim3(256), 0, stream, colours, ids,
                     pixels, encode_ids ? 1 : 0, frame);
  return ok("frame_lookup", err);
}

}  // namespace eye
