// Device-resident scene: the wide BVH, the triangles and their shading data
// in BVH (Morton) order.  Host-side declarations shared by the C ABI, the
// LBVH builder and the render kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "bvh_format.h"
#include "eyecore.h"

namespace eye {

struct DeviceScene {
  WideNode* nodes = nullptr;   // node 0 is the root
  int node_count = 0;
  // Triangles in sorted order, three float4 each:
  //   [3k+0] = (v0.xyz, original triangle id as uint bits)
  //   [3k+1] = (v1.xyz, 0), [3k+2] = (v2.xyz, 0)
  float4* tris = nullptr;
  // Shading data in sorted order, three float4 each:
  //   [3k+0] = (c0.rgb, colour mode word as uint bits), [3k+1] = (c1.rgb, 0), [3k+2] = (c2.rgb, 0)
  float4* cols = nullptr;
  uint32_t* sorted_ids = nullptr;  // sorted position -> original triangle id
  uint32_t* codes = nullptr;       // sorted Morton codes
  int tri_count = 0;
  int stack_bound = 0;             // deepest traversal stack any ray can need
  int levels = 0;
  int leaf_max = 4;
  int sky_mode = 0;
  // Base-colour textures (not owned: the scene loader's upload).  A textured
  // triangle's cols hold (u0, v0, u1, mode) and (v1, u2, v2, 0).
  const uint32_t* texels = nullptr;
  const TexView* textures = nullptr;
  int texture_count = 0;
  double build_ms = 0.0;
  // Distinct for every scene the library builds (a process-wide counter set
  // after each build): caches keyed on a scene compare it, because a rebuilt
  // scene can reuse the freed buffers' addresses.
  uint64_t generation = 0;
  // Device address of a host-visible word (not owned) that a packet traversal
  // entered by a partial wave sets; null: not reported.
  uint32_t* fault = nullptr;
  float centroid_lo[3] = {0, 0, 0}, centroid_hi[3] = {0, 0, 0};
};

// Builds the BVH on the GPU from world-space triangles (9 floats each, original
// order), per-vertex colours (9 floats each) and per-triangle colour modes.
// Device inputs; the result owns its buffers.  Returns false and sets `err`.
bool build_device_scene(const float* d_pos, const float* d_col, const uint32_t* d_mode, int n,
                        int leaf_max, hipStream_t stream, DeviceScene* out, std::string* err);

void free_device_scene(DeviceScene* s);

}  // namespace eye
