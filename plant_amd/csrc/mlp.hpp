// Tree-size predictor (see mlp.hip): constants, parameter layout and launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace chl {

constexpr int kMlpF = 8;   // input features (4 used, zero padded to the MFMA K of 8)
constexpr int kMlpH = 64;  // hidden width of both layers
constexpr int kMlpO = 2;   // outputs: log1p(tree size), log1p(label count)
// FP32 master parameter vector: W1 [F][H], b1 [H], W2 [H][H], b2 [H], W3 [H][O], b3 [O]
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpF * kMlpH;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpH;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpH * kMlpH;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpH;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpH * kMlpO;
constexpr int kMlpParams = kMlpOffB3 + kMlpO;

// Y[N][O] = MLP(X[N][F]) with the BF16 weights wb and FP32 biases from P.
hipError_t mlp_forward(const float* X, int N, const uint16_t* wb, const float* P, float* Y, hipStream_t s);
// G += d/dP of mean over (B rows of idx) x O of (MLP(X[idx]) - T)^2; loss += sum of squared errors.
hipError_t mlp_grad(const float* X, const float* T, const uint32_t* idx, int B, const uint16_t* wb, const float* P,
                    float* G, float* loss, hipStream_t s);
// One Adam step (beta 0.9/0.999); *step (device memory, 0 before the first step)
// is incremented first and used as t. Refreshes the BF16 weights.
hipError_t mlp_adam(float* P, const float* G, float* M, float* V, uint16_t* wb, float lr, int* step, hipStream_t s);
hipError_t mlp_sync(const float* P, uint16_t* wb, hipStream_t s);
// D = A B for one 32x32x8 BF16 tile (row-major operands).
hipError_t mfma_probe(const uint16_t* A, const uint16_t* B, float* D, hipStream_t s);

}  // namespace chl
