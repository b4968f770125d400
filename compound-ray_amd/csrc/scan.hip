#include "scan.h"

namespace eye {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanBlock = kScanThreads * kScanItems;

// Inclusive scan of 256 values held one per thread, via LDS (Hillis-Steele on
// wave totals after an in-wave shuffle scan).
__device__ uint32_t block_inclusive(uint32_t v, uint32_t* lds_waves) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  if (lane == 63) lds_waves[w] = v;
  __syncthreads();
  uint32_t add = 0;
  for (int k = 0; k < w; ++k) add += lds_waves[k];
  __syncthreads();
  return v + add;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(const uint32_t* __restrict__ in,
                                                              uint32_t* __restrict__ out, size_t n,
                                                              uint32_t* __restrict__ sums) {
  __shared__ uint32_t waves[kScanThreads / 64];
  const size_t base = (size_t)blockIdx.x * kScanBlock + (size_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t local = 0;
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    local += v[k];
  }
  const uint32_t incl = block_inclusive(local, waves);
  uint32_t run = incl - local;
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = incl;
}

// Single block, any length: exclusive scan of the block sums, total at the end.
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(uint32_t* __restrict__ sums, size_t n,
                                                            uint32_t* __restrict__ total) {
  __shared__ uint32_t waves[kScanThreads / 64];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (size_t base = 0; base < n; base += kScanThreads) {
    const size_t i = base + threadIdx.x;
    const uint32_t v = i < n ? sums[i] : 0u;
    const uint32_t incl = block_inclusive(v, waves);
    const uint32_t c = carry;
    if (i < n) sums[i] = c + incl - v;
    __syncthreads();
    if (threadIdx.x == kScanThreads - 1) carry = c + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(uint32_t* __restrict__ out, size_t n,
                                                           const uint32_t* __restrict__ sums) {
  const uint32_t add = sums[blockIdx.x];
  const size_t base = (size_t)blockIdx.x * kScanBlock + (size_t)threadIdx.x * kScanItems;
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) out[base + k] += add;
}

}  // namespace

bool scan_reserve(ScanScratch* s, size_t n, std::string* err) {
  const size_t blocks = (n + kScanBlock - 1) / kScanBlock + 1;
  if (s->capacity >= n && s->block_sums) return true;
  scan_release(s);
  hipError_t e = hipMalloc(&s->block_sums, sizeof(uint32_t) * blocks);
  if (e == hipSuccess) e = hipMalloc(&s->total, sizeof(uint32_t));
  if (e == hipSuccess) e = hipHostMalloc(&s->host_total, sizeof(uint32_t), hipHostMallocDefault);
  if (e != hipSuccess) {
    if (err) *err = std::string("scan: allocation failed: ") + hipGetErrorString(e);
    scan_release(s);
    return false;
  }
  s->capacity = n;
  return true;
}

void scan_release(ScanScratch* s) {
  if (s->block_sums) (void)hipFree(s->block_sums);
  if (s->total) (void)hipFree(s->total);
  if (s->host_total) (void)hipHostFree(s->host_total);
  *s = ScanScratch();
}

static bool scan_impl(ScanScratch* s, const uint32_t* in, uint32_t* out, size_t n, hipStream_t stream,
                      std::string* err) {
  if (n > s->capacity) {
    if (err) *err = "scan: input larger than the reserved scratch";
    return false;
  }
  if (n == 0) {
    (void)hipMemsetAsync(s->total, 0, sizeof(uint32_t), stream);
    return true;
  }
  const size_t blocks = (n + kScanBlock - 1) / kScanBlock;
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)blocks), dim3(kScanThreads), 0, stream, in, out, n,
                     s->block_sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, stream, s->block_sums, blocks, s->total);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)blocks), dim3(kScanThreads), 0, stream, out, n, s->block_sums);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    if (err) *err = std::string("scan: launch failed: ") + hipGetErrorString(e);
    return false;
  }
  return true;
}

bool scan_exclusive_u32(ScanScratch* s, uint32_t* data, size_t n, hipStream_t stream, std::string* err) {
  return scan_impl(s, data, data, n, stream, err);
}

bool scan_exclusive_u32_total(ScanScratch* s, const uint32_t* in, uint32_t* out, size_t n,
                              hipStream_t stream, uint32_t* total, std::string* err) {
  if (!scan_impl(s, in, out, n, stream, err)) return false;
  hipError_t e = hipMemcpyAsync(s->host_total, s->total, sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    if (err) *err = std::string("scan: total readback failed: ") + hipGetErrorString(e);
    return false;
  }
  *total = *s->host_total;
  return true;
}

}  // namespace eye
