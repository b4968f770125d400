// Fixed device scratch for the allocation-free batch render
// (renderPoseBatchInto): callers under HIP graph capture or in tight serving
// loops cannot let the library call hipMalloc, so the work counters of that
// call's unit renderer live in a file-scope device array instead of a
// library allocation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "render.h"

namespace eye {

// Words of the fixed scratch array (256 bytes per device, held until the
// process ends); the unit renderer's queue counters fit in it.
constexpr int kNoAllocScratchWords = 64;
static_assert(kNoAllocScratchWords >= kQueueWords, "the queue counters must fit the fixed scratch");

// Largest (pose, ommatidium) unit count of one allocation-free call (the unit
// renderer's queue positions are 32-bit).
constexpr uint64_t kNoAllocMaxUnits = 0xffffffffull - 65536ull;

// Looks up `device`'s copy of the scratch array (once per device, with that
// device current).  Returns false and sets err on failure.
bool no_alloc_scratch_init(int device, std::string* err);

// The device address of `device`'s scratch array, or null before
// no_alloc_scratch_init succeeded for it.
uint32_t* no_alloc_scratch(int device);

}  // namespace eye
