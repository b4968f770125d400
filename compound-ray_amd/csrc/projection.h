// Spherical projections of compound-eye renders: the window shows, per pixel,
// the ommatidium closest in angle to the pixel's unit-sphere direction, by the
// ommatidia's positions (spherical_positionwise[_ids]), axes
// (spherical_orientationwise[_ids]) or axes restricted to the pixel's half of
// the eye (spherical_split_orientationwise), shaders.cu:412-640.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "eyecore.h"

namespace eye {

// ids[y * width + x] for every pixel of the window (computed once per eye and
// window size).
bool launch_spherical_ids(const Ommatidium* omms, int omm_count, int width, int height, int variant, uint32_t* ids,
                          hipStream_t stream, std::string* err);

// frame[p] = make_color(colours[ids[p]]) with alpha 255, or (encode_ids) the
// ommatidium index in the bytes (r, g, b, a) = (id >> 24, id >> 16, id >> 8, id).
bool launch_frame_lookup(const float4* colours, const uint32_t* ids, int pixels, bool encode_ids, uchar4* frame,
                         hipStream_t stream, std::string* err);

}  // namespace eye
