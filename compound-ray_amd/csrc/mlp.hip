// BF16 MFMA kernels of the learned sampler MLP (see mlp.h for the rounding
// points).  Activations are kept transposed (features x samples) so that the
// accumulator layout of one v_mfma_f32_16x16x16bf16 is directly the B operand
// of the next layer: a wave pushes 16 samples through all four layers without
// moving data between lanes.  The 64-input products pair two such blocks into
// one gfx950 v_mfma_f32_16x16x32_bf16.  Weight gradients need the samples as the
// reduction dimension, which one LDS transpose per layer provides.
#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

namespace eye {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kGradBlocks = 64;

// packed BF16 layouts (uint16 units)
constexpr int P_W1T = 0;                 // [64 out][16 in]
constexpr int P_W2T = P_W1T + 64 * 16;   // [64 out][64 in]
constexpr int P_W3T = P_W2T + 64 * 64;
constexpr int P_W2N = P_W3T + 64 * 64;   // [64 in][64 out]
constexpr int P_W3N = P_W2N + 64 * 64;
constexpr int P_W4 = P_W3N + 64 * 64;    // [64]
constexpr int P_TOTAL = P_W4 + 64;       // 17472 (multiple of 8)
static_assert(P_TOTAL % 8 == 0, "packed size must allow 16-byte copies");

// FP32 -> BF16, round to nearest even (a NaN stays a quiet NaN).  On the GPU
// this is gfx950's one-instruction conversion (v_cvt_pk_bf16_f32).
__host__ __device__ inline uint16_t f2bf(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_bit_cast(uint16_t, (__bf16)f);
#else
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
#endif
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ inline float bf2f_s(short b) { return bf2f((uint16_t)b); }

__device__ inline f4 mfma(s4 a, s4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// Two K = 16 products as one gfx950 K = 32 MFMA.  Lane group g = l/16 gives
// the instruction k slots 8g .. 8g + 7; here slots 8g .. 8g + 3 hold the
// first block's k = 4g .. 4g + 3 and slots 8g + 4 .. 8g + 7 the second's.  The
// sum runs over every slot, so A and B only have to share that order.
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
__device__ inline f4 mfma2(s4 a0, s4 a1, s4 b0, s4 b1, f4 c) {
  const s8 a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
  const s8 b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a), __builtin_bit_cast(b8, b), c, 0, 0, 0);
}

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ inline s4 lds4(const uint16_t* p) { return *reinterpret_cast<const s4*>(p); }

__global__ void k_pack(const float* __restrict__ m, uint16_t* __restrict__ p) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 64 * 16) {
    const int o = t / 16, i = t % 16;
    p[P_W1T + t] = i < kMlpIn ? f2bf(m[MlpOffsets::W1 + i * 64 + o]) : (uint16_t)0;
  }
  if (t < 64 * 64) {
    const int o = t / 64, i = t % 64;
    p[P_W2T + t] = f2bf(m[MlpOffsets::W2 + i * 64 + o]);
    p[P_W3T + t] = f2bf(m[MlpOffsets::W3 + i * 64 + o]);
    p[P_W2N + t] = f2bf(m[MlpOffsets::W2 + t]);
    p[P_W3N + t] = f2bf(m[MlpOffsets::W3 + t]);
  }
  if (t < 64) p[P_W4 + t] = f2bf(m[MlpOffsets::W4 + t]);
}

// Copies the packed weights and the biases into LDS (whole block).
__device__ inline void load_weights(const uint16_t* __restrict__ packed, const float* __restrict__ master,
                                    uint16_t* w, float* bias) {
  const uint4* src = reinterpret_cast<const uint4*>(packed);
  uint4* dst = reinterpret_cast<uint4*>(w);
  for (int k = threadIdx.x; k < P_TOTAL / 8; k += blockDim.x) dst[k] = src[k];
  for (int k = threadIdx.x; k < 64; k += blockDim.x) {
    bias[k] = master[MlpOffsets::b1 + k];
    bias[64 + k] = master[MlpOffsets::b2 + k];
    bias[128 + k] = master[MlpOffsets::b3 + k];
  }
  if (threadIdx.x == 0) bias[192] = master[MlpOffsets::b4];
}

// Input tile as the B operand of layer 1: lane holds X[s0 + l%16][4 (l/16) + j].
__device__ inline s4 load_x(const float* __restrict__ x, int64_t n, int64_t s0) {
  const int l = threadIdx.x & 63, q = l >> 4;
  const int64_t s = s0 + (l & 15);
  s4 v;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int f = 4 * q + j;
    v[j] = (s < n && f < kMlpIn) ? (short)f2bf(x[s * kMlpIn + f]) : (short)0;
  }
  return v;
}

struct Acts {
  s4 h1[4], h2[4], h3[4];
  float y;
};

// One hidden layer, transposed: H_out^T = relu(W^T H_in^T + b), `nk` input blocks.
template <int kIn, int kOff>
__device__ inline void hidden_layer(const uint16_t* w, const float* b, const s4* in, s4* out) {
  const int l = threadIdx.x & 63, l16 = l & 15, q = l >> 4;
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const uint16_t* row = w + kOff + (ob * 16 + l16) * kIn + 4 * q;
    if constexpr (kIn % 32 == 0) {
#pragma unroll
      for (int kb = 0; kb < kIn / 16; kb += 2)
        acc = mfma2(lds4(row + kb * 16), lds4(row + kb * 16 + 16), in[kb], in[kb + 1], acc);
    } else {
#pragma unroll
      for (int kb = 0; kb < kIn / 16; ++kb) acc = mfma(lds4(row + kb * 16), in[kb], acc);
    }
    s4 h;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float z = acc[i] + b[ob * 16 + 4 * q + i];
      h[i] = (short)f2bf(z > 0.0f ? z : 0.0f);
    }
    out[ob] = h;
  }
}

__device__ inline void forward_tile(const uint16_t* w, const float* bias, s4 xb, Acts& a) {
  const int l = threadIdx.x & 63, q = l >> 4;
  hidden_layer<16, P_W1T>(w, bias, &xb, a.h1);
  hidden_layer<64, P_W2T>(w, bias + 64, a.h1, a.h2);
  hidden_layer<64, P_W3T>(w, bias + 128, a.h2, a.h3);
  float part = 0.0f;
#pragma unroll
  for (int ob = 0; ob < 4; ++ob)
#pragma unroll
    for (int i = 0; i < 4; ++i) part += bf2f_s(a.h3[ob][i]) * bf2f(w[P_W4 + ob * 16 + 4 * q + i]);
  part += __shfl_xor(part, 16);
  part += __shfl_xor(part, 32);
  a.y = part + bias[192];
}

__global__ __launch_bounds__(kThreads) void k_forward(const float* __restrict__ x, int64_t n,
                                                      const uint16_t* __restrict__ packed,
                                                      const float* __restrict__ master, float* __restrict__ y) {
  __shared__ __align__(16) uint16_t w[P_TOTAL];
  __shared__ float bias[193];
  load_weights(packed, master, w, bias);
  __syncthreads();
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t tiles = (n + 15) / 16;
  for (int64_t t = (int64_t)blockIdx.x * kWaves + wave; t < tiles; t += (int64_t)gridDim.x * kWaves) {
    Acts a;
    forward_tile(w, bias, load_x(x, n, t * 16), a);
    const int64_t s = t * 16 + l;
    if (l < 16 && s < n) y[s] = a.y;
  }
}

// Writes a 16x16 BF16 tile held in accumulator layout (lane: rows 4 (l/16) + i,
// column l%16) to LDS as [row][column].
__device__ inline void store_tile(uint16_t* tile, s4 v) {
  const int l = threadIdx.x & 63, l16 = l & 15, q = l >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i) tile[(4 * q + i) * 16 + l16] = (uint16_t)v[i];
}
// Reads row l%16, columns 4 (l/16) .. +3 of an LDS tile: the transposed
// operand (A with k = samples, or B with k = samples).
__device__ inline s4 load_tile_t(const uint16_t* tile) {
  const int l = threadIdx.x & 63, l16 = l & 15, q = l >> 4;
  return lds4(tile + l16 * 16 + 4 * q);
}

// Back-propagates through one hidden layer given dz of its output (fp32 and
// bf16, accumulator layout) and its input activations; accumulates dW into
// dw[in block][out block] (accumulator layout: row = in, column = out).
template <int kInBlocks>
__device__ inline void weight_grad(uint16_t* tiles_h, uint16_t* tiles_d, const s4* h_in, const s4* dzb,
                                   f4 (*dw)[4]) {
  for (int b = 0; b < kInBlocks; ++b) store_tile(tiles_h + b * 256, h_in[b]);
  for (int b = 0; b < 4; ++b) store_tile(tiles_d + b * 256, dzb[b]);
  wave_sync();
  s4 bt[4];
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) bt[ob] = load_tile_t(tiles_d + ob * 256);
#pragma unroll
  for (int ib = 0; ib < kInBlocks; ++ib) {
    const s4 at = load_tile_t(tiles_h + ib * 256);
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) dw[ib][ob] = mfma(at, bt[ob], dw[ib][ob]);
  }
  wave_sync();
}

// dh_in^T = W dz^T for a 64-in layer (W in-major in LDS), then the ReLU mask of
// h_in: returns dz of the input layer in fp32 and bf16.
__device__ inline void back_layer(const uint16_t* wn, const s4* dzb, const s4* h_in, float (*dz_out)[4],
                                  s4* dzb_out) {
  const int l = threadIdx.x & 63, l16 = l & 15, q = l >> 4;
#pragma unroll
  for (int ib = 0; ib < 4; ++ib) {
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const uint16_t* row = wn + (ib * 16 + l16) * 64 + 4 * q;
#pragma unroll
    for (int kb = 0; kb < 4; kb += 2) acc = mfma2(lds4(row + kb * 16), lds4(row + kb * 16 + 16), dzb[kb], dzb[kb + 1], acc);
    s4 o;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float d = (h_in[ib][i] > 0) ? acc[i] : 0.0f;   // bf16 bits of a non-negative value: > 0 iff positive
      dz_out[ib][i] = d;
      o[i] = (short)f2bf(d);
    }
    dzb_out[ib] = o;
  }
}

__device__ inline float sum16(float v) {  // over the 16 lanes sharing l/16
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 4);
  v += __shfl_xor(v, 8);
  return v;
}

__global__ __launch_bounds__(kThreads) void k_grad(const float* __restrict__ x, const float* __restrict__ target,
                                                   int64_t n, const uint16_t* __restrict__ packed,
                                                   const float* __restrict__ master, float* __restrict__ partial,
                                                   float* __restrict__ loss_partial) {
  __shared__ __align__(16) uint16_t w[P_TOTAL];
  __shared__ float bias[193];
  __shared__ __align__(16) uint16_t tiles[kWaves][2][4 * 256];
  load_weights(packed, master, w, bias);
  __syncthreads();
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, l16 = l & 15, q = l >> 4;
  uint16_t* th = tiles[wave][0];
  uint16_t* td = tiles[wave][1];
  f4 dw1[1][4], dw2[4][4], dw3[4][4];
  float dw4[4][4], db1[4][4], db2[4][4], db3[4][4];
  float db4 = 0.0f, loss = 0.0f;
  const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int a = 0; a < 4; ++a) {
    dw1[0][a] = zero;
    for (int b = 0; b < 4; ++b) {
      dw2[a][b] = zero;
      dw3[a][b] = zero;
      dw4[a][b] = db1[a][b] = db2[a][b] = db3[a][b] = 0.0f;
    }
  }
  const float inv_n = 1.0f / (float)n;
  const int64_t tiles_n = (n + 15) / 16;
  for (int64_t t = (int64_t)blockIdx.x * kWaves + wave; t < tiles_n; t += (int64_t)gridDim.x * kWaves) {
    const int64_t s0 = t * 16;
    const s4 xb = load_x(x, n, s0);
    Acts a;
    forward_tile(w, bias, xb, a);
    const int64_t s = s0 + l16;
    const bool valid = s < n;
    const float err = valid ? a.y - target[s] : 0.0f;
    const float dy = 2.0f * err * inv_n;
    if (q == 0) {
      loss += err * err;
      db4 += dy;
    }
    // output layer: dW4 and dz3
    float dz3[4][4];
    s4 dz3b[4];
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
      s4 o;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float h = bf2f_s(a.h3[ob][i]);
        dw4[ob][i] += dy * h;
        const float d = (a.h3[ob][i] > 0) ? bf2f(w[P_W4 + ob * 16 + 4 * q + i]) * dy : 0.0f;
        dz3[ob][i] = d;
        db3[ob][i] += d;
        o[i] = (short)f2bf(d);
      }
      dz3b[ob] = o;
    }
    weight_grad<4>(th, td, a.h2, dz3b, dw3);
    float dz2[4][4];
    s4 dz2b[4];
    back_layer(w + P_W3N, dz3b, a.h2, dz2, dz2b);
    for (int ob = 0; ob < 4; ++ob)
      for (int i = 0; i < 4; ++i) db2[ob][i] += dz2[ob][i];
    weight_grad<4>(th, td, a.h1, dz2b, dw2);
    float dz1[4][4];
    s4 dz1b[4];
    back_layer(w + P_W2N, dz2b, a.h1, dz1, dz1b);
    for (int ob = 0; ob < 4; ++ob)
      for (int i = 0; i < 4; ++i) db1[ob][i] += dz1[ob][i];
    weight_grad<1>(th, td, &xb, dz1b, dw1);
  }
  // deterministic cross-wave sum: wave 0 writes, then waves 1..3 add in order
  float* out = partial + (size_t)blockIdx.x * kMlpParams;
  for (int wv = 0; wv < kWaves; ++wv) {
    if (wave == wv) {
      auto put = [&](int idx, float v) {
        if (wv == 0) out[idx] = v;
        else out[idx] += v;
      };
      for (int ob = 0; ob < 4; ++ob)
        for (int i = 0; i < 4; ++i) {
          const int row = 4 * q + i, col = ob * 16 + l16;
          if (row < kMlpIn) put(MlpOffsets::W1 + row * 64 + col, dw1[0][ob][i]);
          for (int ib = 0; ib < 4; ++ib) {
            put(MlpOffsets::W2 + (ib * 16 + row) * 64 + col, dw2[ib][ob][i]);
            put(MlpOffsets::W3 + (ib * 16 + row) * 64 + col, dw3[ib][ob][i]);
          }
          const float v4 = sum16(dw4[ob][i]);
          const float v1 = sum16(db1[ob][i]);
          const float v2 = sum16(db2[ob][i]);
          const float v3 = sum16(db3[ob][i]);
          if (l16 == 0) {
            const int f = ob * 16 + 4 * q + i;
            put(MlpOffsets::W4 + f, v4);
            put(MlpOffsets::b1 + f, v1);
            put(MlpOffsets::b2 + f, v2);
            put(MlpOffsets::b3 + f, v3);
          }
        }
      const float vb4 = sum16(db4);
      const float vl = sum16(loss);
      if (l == 0) {
        put(MlpOffsets::b4, vb4);
        if (wv == 0) loss_partial[blockIdx.x] = vl;
        else loss_partial[blockIdx.x] += vl;
      }
    }
    __syncthreads();
  }
}

// grad[p] = sum over blocks in block order; loss = sum / n.
__global__ void k_reduce_grad(const float* __restrict__ partial, const float* __restrict__ loss_partial, int blocks,
                              int64_t n, float* __restrict__ grad, float* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kMlpParams) {
    float s = 0.0f;
    for (int b = 0; b < blocks; ++b) s += partial[(size_t)b * kMlpParams + p];
    grad[p] = s;
  }
  if (p == 0) {
    float s = 0.0f;
    for (int b = 0; b < blocks; ++b) s += loss_partial[b];
    *loss = s / (float)n;
  }
}

__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, float lr, float bc1, float bc2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  const float gi = g[p];
  const float mi = 0.9f * m[p] + 0.1f * gi;
  const float vi = 0.999f * v[p] + 0.001f * gi * gi;
  m[p] = mi;
  v[p] = vi;
  w[p] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
}

__global__ void k_gather_rows(const float* __restrict__ x, const float* __restrict__ t,
                              const uint32_t* __restrict__ idx, int64_t n, float* __restrict__ xo,
                              float* __restrict__ to) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint32_t s = idx[r];
  for (int f = 0; f < kMlpIn; ++f) xo[r * kMlpIn + f] = x[(int64_t)s * kMlpIn + f];
  to[r] = t[s];
}

bool launch_ok(const char* what, std::string* err) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    if (err) *err = std::string("mlp: ") + what + ": " + hipGetErrorString(e);
    return false;
  }
  return true;
}

template <typename T>
bool alloc(T** p, size_t n, std::string* err) {
  const hipError_t e = hipMalloc(p, sizeof(T) * n);
  if (e != hipSuccess) {
    if (err) *err = std::string("mlp: hipMalloc: ") + hipGetErrorString(e);
    return false;
  }
  return true;
}

}  // namespace

void Mlp::release() {
  for (void* p : {(void*)master_, (void*)adam_m_, (void*)adam_v_, (void*)packed_, (void*)partial_,
                  (void*)loss_partial_, (void*)grad_, (void*)loss_, (void*)xb_, (void*)tb_, (void*)perm_})
    if (p) (void)hipFree(p);
  master_ = adam_m_ = adam_v_ = partial_ = loss_partial_ = grad_ = loss_ = xb_ = tb_ = nullptr;
  packed_ = nullptr;
  perm_ = nullptr;
  batch_cap_ = perm_cap_ = 0;
  step_ = 0;
}

bool Mlp::pack(hipStream_t stream, std::string* err) {
  hipLaunchKernelGGL(k_pack, dim3(16), dim3(256), 0, stream, master_, packed_);
  return launch_ok("pack", err);
}

bool Mlp::init(uint32_t seed, hipStream_t stream, std::string* err) {
  if (!master_) {
    if (!alloc(&master_, kMlpParams, err) || !alloc(&adam_m_, kMlpParams, err) ||
        !alloc(&adam_v_, kMlpParams, err) || !alloc(&packed_, P_TOTAL, err) ||
        !alloc(&partial_, (size_t)kGradBlocks * kMlpParams, err) || !alloc(&loss_partial_, kGradBlocks, err) ||
        !alloc(&grad_, kMlpParams, err) || !alloc(&loss_, 1, err))
      return false;
  }
  // He-uniform initialisation from a seeded host generator, zero biases
  std::vector<float> h(kMlpParams, 0.0f);
  std::mt19937 gen(seed);
  auto fill = [&](int off, int fan_in, int count) {
    const float lim = std::sqrt(6.0f / (float)fan_in);
    std::uniform_real_distribution<float> u(-lim, lim);
    for (int k = 0; k < count; ++k) h[(size_t)(off + k)] = u(gen);
  };
  fill(MlpOffsets::W1, kMlpIn, kMlpIn * 64);
  fill(MlpOffsets::W2, 64, 64 * 64);
  fill(MlpOffsets::W3, 64, 64 * 64);
  fill(MlpOffsets::W4, 64, 64);
  step_ = 0;
  if (!set_params(h.data(), stream, err)) return false;
  return true;
}

bool Mlp::set_params(const float* host, hipStream_t stream, std::string* err) {
  if (!master_ && !init(0, stream, err)) return false;
  hipError_t e = hipMemcpyAsync(master_, host, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipMemsetAsync(adam_m_, 0, sizeof(float) * kMlpParams, stream);
  if (e == hipSuccess) e = hipMemsetAsync(adam_v_, 0, sizeof(float) * kMlpParams, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    if (err) *err = std::string("mlp: set_params: ") + hipGetErrorString(e);
    return false;
  }
  step_ = 0;
  return pack(stream, err);
}

bool Mlp::get_params(float* host, hipStream_t stream, std::string* err) {
  if (!master_) {
    if (err) *err = "mlp: not initialised";
    return false;
  }
  hipError_t e = hipMemcpyAsync(host, master_, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    if (err) *err = std::string("mlp: get_params: ") + hipGetErrorString(e);
    return false;
  }
  return true;
}

bool Mlp::predict(const float* d_x, int64_t n, float* d_y, hipStream_t stream, std::string* err) {
  if (!master_) {
    if (err) *err = "mlp: not initialised";
    return false;
  }
  if (n <= 0) return true;
  const int64_t tiles = (n + 15) / 16;
  const int64_t blocks = std::min<int64_t>((tiles + kWaves - 1) / kWaves, 4096);
  hipLaunchKernelGGL(k_forward, dim3((unsigned)blocks), dim3(kThreads), 0, stream, d_x, n, packed_, master_, d_y);
  return launch_ok("forward", err);
}

bool Mlp::loss_grad(const float* d_x, const float* d_t, int64_t n, float* d_loss, float* d_grad,
                    hipStream_t stream, std::string* err) {
  if (!master_) {
    if (err) *err = "mlp: not initialised";
    return false;
  }
  if (n <= 0) {
    if (err) *err = "mlp: empty batch";
    return false;
  }
  const int64_t tiles = (n + 15) / 16;
  const int blocks = (int)std::min<int64_t>(kGradBlocks, (tiles + kWaves - 1) / kWaves);
  hipLaunchKernelGGL(k_grad, dim3(blocks), dim3(kThreads), 0, stream, d_x, d_t, n, packed_, master_, partial_,
                     loss_partial_);
  if (!launch_ok("grad", err)) return false;
  hipLaunchKernelGGL(k_reduce_grad, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream, partial_, loss_partial_,
                     blocks, n, d_grad, d_loss);
  return launch_ok("reduce_grad", err);
}

bool Mlp::adam_step(const float* d_grad, float lr, hipStream_t stream, std::string* err) {
  ++step_;
  const float bc1 = 1.0f - std::pow(0.9f, (float)step_);
  const float bc2 = 1.0f - std::pow(0.999f, (float)step_);
  hipLaunchKernelGGL(k_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream, master_, adam_m_, adam_v_, d_grad,
                     lr, bc1, bc2);
  if (!launch_ok("adam", err)) return false;
  return pack(stream, err);
}

bool Mlp::train(const float* d_x, const float* d_t, int64_t n, int epochs, int batch, float lr, uint32_t seed,
                std::vector<float>* losses, hipStream_t stream, std::string* err) {
  if (!master_ && !init(seed, stream, err)) return false;
  if (n <= 0 || batch <= 0) {
    if (err) *err = "mlp: empty training set";
    return false;
  }
  if (batch_cap_ < batch) {
    if (xb_) (void)hipFree(xb_);
    if (tb_) (void)hipFree(tb_);
    xb_ = tb_ = nullptr;
    if (!alloc(&xb_, (size_t)batch * kMlpIn, err) || !alloc(&tb_, (size_t)batch, err)) return false;
    batch_cap_ = batch;
  }
  if (perm_cap_ < n) {
    if (perm_) (void)hipFree(perm_);
    perm_ = nullptr;
    if (!alloc(&perm_, (size_t)n, err)) return false;
    perm_cap_ = n;
  }
  std::vector<uint32_t> perm((size_t)n);
  std::mt19937 gen(seed);
  std::vector<float> batch_losses;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int64_t k = 0; k < n; ++k) perm[(size_t)k] = (uint32_t)k;
    std::shuffle(perm.begin(), perm.end(), gen);
    if (hipMemcpyAsync(perm_, perm.data(), sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice, stream) !=
        hipSuccess) {
      if (err) *err = "mlp: permutation upload failed";
      return false;
    }
    double sum = 0.0;
    int64_t count = 0;
    std::vector<float> lb;
    for (int64_t b0 = 0; b0 < n; b0 += batch) {
      const int64_t bn = std::min<int64_t>(batch, n - b0);
      hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((bn + 255) / 256)), dim3(256), 0, stream, d_x, d_t,
                         perm_ + b0, bn, xb_, tb_);
      if (!launch_ok("gather", err)) return false;
      if (!loss_grad(xb_, tb_, bn, loss_, grad_, stream, err)) return false;
      if (!adam_step(grad_, lr, stream, err)) return false;
      float l = 0.0f;
      if (hipMemcpyAsync(&l, loss_, sizeof(float), hipMemcpyDeviceToHost, stream) != hipSuccess ||
          hipStreamSynchronize(stream) != hipSuccess) {
        if (err) *err = "mlp: loss readback failed";
        return false;
      }
      sum += (double)l * (double)bn;
      count += bn;
    }
    if (losses) losses->push_back((float)(sum / (double)count));
  }
  return true;
}

}  // namespace eye
