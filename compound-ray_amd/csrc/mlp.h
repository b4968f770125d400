// Learned adaptive sampler: a 16 -> 64 -> 64 -> 64 -> 1 ReLU MLP in BF16 on
// MFMA with FP32 accumulation, with hand-written forward, backward and Adam
// kernels for gfx950.  No reference counterpart (the reference samples every
// ommatidium uniformly, libEyeRenderer.cpp:167-176).
//
// Parameters (FP32 master copy, this order): W1[16][64], b1[64], W2[64][64],
// b2[64], W3[64][64], b3[64], W4[64], b4 — W[in][out].
// Rounding points (the oracle mirrors them): inputs, weights, hidden
// activations and back-propagated pre-activation gradients are rounded to
// BF16 (round to nearest even) where they enter an MFMA; products accumulate
// in FP32; biases, the output layer's sum and all gradients are FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace eye {

constexpr int kMlpIn = 16;
constexpr int kMlpHidden = 64;
constexpr int kMlpParams = kMlpIn * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1;  // 9473

struct MlpOffsets {
  static constexpr int W1 = 0;
  static constexpr int b1 = W1 + kMlpIn * 64;
  static constexpr int W2 = b1 + 64;
  static constexpr int b2 = W2 + 64 * 64;
  static constexpr int W3 = b2 + 64;
  static constexpr int b3 = W3 + 64 * 64;
  static constexpr int W4 = b3 + 64;
  static constexpr int b4 = W4 + 64;
};

class Mlp {
 public:
  ~Mlp() { release(); }
  bool init(uint32_t seed, hipStream_t stream, std::string* err);
  bool set_params(const float* host, hipStream_t stream, std::string* err);
  bool get_params(float* host, hipStream_t stream, std::string* err);
  // y[n] = f(X[n][kMlpIn]); device pointers.
  bool predict(const float* d_x, int64_t n, float* d_y, hipStream_t stream, std::string* err);
  // Mean squared error and its gradient (FP32, kMlpParams) over rows
  // [0, n) of device arrays; *loss and grad are device outputs.
  bool loss_grad(const float* d_x, const float* d_t, int64_t n, float* d_loss, float* d_grad,
                 hipStream_t stream, std::string* err);
  // One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with gradient d_grad.
  bool adam_step(const float* d_grad, float lr, hipStream_t stream, std::string* err);
  // Minibatch training over device arrays, rows shuffled per epoch with a
  // seeded permutation; losses receives the mean loss of every epoch.
  bool train(const float* d_x, const float* d_t, int64_t n, int epochs, int batch, float lr, uint32_t seed,
             std::vector<float>* losses, hipStream_t stream, std::string* err);
  bool ready() const { return master_ != nullptr; }
  void release();

 private:
  bool pack(hipStream_t stream, std::string* err);
  float* master_ = nullptr;      // kMlpParams
  float* adam_m_ = nullptr;
  float* adam_v_ = nullptr;
  uint16_t* packed_ = nullptr;   // BF16 layouts for the kernels
  float* partial_ = nullptr;     // per-block gradient partials
  float* loss_partial_ = nullptr;
  float* grad_ = nullptr;
  float* loss_ = nullptr;
  float* xb_ = nullptr;          // gathered minibatch
  float* tb_ = nullptr;
  int64_t batch_cap_ = 0;
  uint32_t* perm_ = nullptr;
  int64_t perm_cap_ = 0;
  int64_t step_ = 0;
};

}  // namespace eye
