// Device-wide exclusive prefix sum of uint32 values (three-phase: per-block
// scans, a scan of the block totals, a uniform add).  Used by the radix sort
// and the BVH collapse.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

namespace eye {

struct ScanScratch {
  uint32_t* block_sums = nullptr;
  uint32_t* total = nullptr;   // one value, device
  uint32_t* host_total = nullptr;  // pinned, one value
  size_t capacity = 0;         // elements the scratch can scan
};

bool scan_reserve(ScanScratch* s, size_t n, std::string* err);
void scan_release(ScanScratch* s);
// In place: data[i] <- sum of data[0..i-1].
bool scan_exclusive_u32(ScanScratch* s, uint32_t* data, size_t n, hipStream_t stream, std::string* err);
// out[i] <- sum of in[0..i-1]; *total <- sum of all (synchronises the stream).
bool scan_exclusive_u32_total(ScanScratch* s, const uint32_t* in, uint32_t* out, size_t n,
                              hipStream_t stream, uint32_t* total, std::string* err);

}  // namespace eye
