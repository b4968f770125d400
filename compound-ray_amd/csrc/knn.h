// Nearest ommatidia of every ommatidium of an eye (host code).  The learned
// sampler pools the pilot statistics of these neighbours: adjacent ommatidia
// view adjacent parts of the scene, so their pooled spread is a far less noisy
// sign of an ommatidium's colour variance than its own eight pilot samples.
// No reference counterpart.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace eye {

constexpr int kNeighbours = 6;
// Positions enter the distance divided by this length (scene units), so that
// ommatidia of different eyes merged into one list stay apart while, within an
// eye (millimetres across), the axis directions decide.
constexpr double kNeighbourPosScale = 0.01;

// out[i * k + m], m = 0 .. k-1: the k ommatidia j != i nearest to i under
// d^2 = |a_i - a_j|^2 + |p_i - p_j|^2 / pos_scale^2 (a the unit axis, p the
// position; FP64), in increasing (d^2, j) order.  Ommatidia with fewer than k
// others repeat their own index in the remaining slots.  `omm` holds n records
// of 8 floats (position, direction, acceptance angle, focal offset).
// Cost: O(n log n) for eyes whose ommatidia are distinct; a cluster of m
// exactly coincident ommatidia cannot be split by distance, and every query
// inside it scans the whole cluster, O(m^2) in total.
void nearest_ommatidia(const float* omm, size_t n, int k, double pos_scale, uint32_t* out);

}  // namespace eye
