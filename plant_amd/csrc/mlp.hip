// Tree-size predictor: a 8 -> 64 -> 64 -> 2 multilayer perceptron (ReLU) with
// BF16 weights and FP32 accumulation, inference and training on the matrix
// cores of gfx950 (v_mfma_f32_32x32x8_bf16). It maps per-root features (rank,
// degree, 2-hop degree sum, betweenness score; padded to 8) to log1p(tree size)
// and log1p(label count) and is trained online on the sizes each finished batch
// measures. No counterpart in the reference; the quantities it predicts are the
// (label_count, tree_size) pair that distributed_memory/akiba.hpp:99-103 returns
// and :291-297 turns into the PLaNT switch ratio.
//
// MFMA operand layout (32x32x8, one wavefront): lane l holds A[l % 32][4 (l / 32)
// + i] and B[4 (l / 32) + i][l % 32] for i = 0..3; accumulator register v of lane l
// is D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].
#include "dev.hpp"
#include "mlp.hpp"

namespace chl {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (inputs are finite)
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f32x16 mfma(bf16x4 a, bf16x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int drow(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }

struct WaveLds {
  uint16_t h1[32][kMlpH + 2];  // bf16 activations of layer 1
  float h2[32][kMlpH + 1];     // fp32 activations of layer 2 (post ReLU)
  float dh[32][kMlpH + 1];     // fp32 gradient scratch
  uint16_t dhb[32][kMlpH + 2]; // bf16 copy of dh (MFMA operand)
  float x[32][kMlpF];
  float dy[32][kMlpO];
};

// Forward pass of 32 rows; leaves h1/h2 in LDS and returns y for (row = lane % 32,
// o = lane / 32).
__device__ float forward_tile(const float* X, int r0, int N, const uint16_t* wb, const float* P, WaveLds& s) {
  const int lane = threadIdx.x & 63;
  const int row = lane & 31, kb = lane >> 5;
  // inputs (rows past N are zero)
  for (int e = lane; e < 32 * kMlpF; e += 64) {
    const int r = e / kMlpF, k = e % kMlpF;
    s.x[r][k] = (r0 + r < N) ? X[(size_t)(r0 + r) * kMlpF + k] : 0.0f;
  }
  __syncthreads();
  // layer 1: K = 8
  bf16x4 a;
  for (int i = 0; i < 4; ++i) a[i] = (short)to_bf16(s.x[row][4 * kb + i]);
  for (int t = 0; t < 2; ++t) {
    bf16x4 b;
    for (int i = 0; i < 4; ++i) b[i] = (short)wb[kMlpOffW1 + (4 * kb + i) * kMlpH + t * 32 + row];
    f32x16 acc = {};
    acc = mfma(a, b, acc);
    for (int v = 0; v < 16; ++v) {
      const int c = t * 32 + row;
      const float h = fmaxf(acc[v] + P[kMlpOffB1 + c], 0.0f);
      s.h1[drow(v, lane)][c] = to_bf16(h);
    }
  }
  __syncthreads();
  // layer 2: K = 64 in 8 steps
  f32x16 acc2[2] = {{}, {}};
  for (int st = 0; st < 8; ++st) {
    bf16x4 a2;
    for (int i = 0; i < 4; ++i) a2[i] = (short)s.h1[row][8 * st + 4 * kb + i];
    for (int t = 0; t < 2; ++t) {
      bf16x4 b;
      for (int i = 0; i < 4; ++i) b[i] = (short)wb[kMlpOffW2 + (8 * st + 4 * kb + i) * kMlpH + t * 32 + row];
      acc2[t] = mfma(a2, b, acc2[t]);
    }
  }
  for (int t = 0; t < 2; ++t)
    for (int v = 0; v < 16; ++v) {
      const int c = t * 32 + row;
      s.h2[drow(v, lane)][c] = fmaxf(acc2[t][v] + P[kMlpOffB2 + c], 0.0f);
    }
  __syncthreads();
  // layer 3: N = 2, FP32 dot products over BF16 weights
  float y = P[kMlpOffB3 + kb];
  for (int c = 0; c < kMlpH; ++c) y += s.h2[row][c] * from_bf16(wb[kMlpOffW3 + c * kMlpO + kb]);
  return y;
}

__global__ __launch_bounds__(64) void mlp_forward_kernel(const float* X, int N, const uint16_t* wb, const float* P,
                                                         float* Y) {
  __shared__ WaveLds s;
  const int lane = threadIdx.x;
  for (int r0 = blockIdx.x * 32; r0 < N; r0 += gridDim.x * 32) {
    const float y = forward_tile(X, r0, N, wb, P, s);
    const int r = r0 + (lane & 31);
    if (r < N) Y[(size_t)r * kMlpO + (lane >> 5)] = y;
    __syncthreads();
  }
}

// Gradient of the mean squared error over a minibatch (rows idx[0..B) of X and
// T), summed into G (FP32, zeroed by the caller). scale = 2 / (B * kMlpO).
__global__ __launch_bounds__(64) void mlp_grad_kernel(const float* X, const float* T, const uint32_t* idx, int B,
                                                      const uint16_t* wb, const float* P, float* G, float scale,
                                                      float* loss) {
  __shared__ WaveLds s;
  __shared__ float xs[32 * kMlpF];
  const int lane = threadIdx.x;
  const int row = lane & 31, kb = lane >> 5;
  for (int r0 = blockIdx.x * 32; r0 < B; r0 += gridDim.x * 32) {
    // gather the minibatch rows into a dense tile
    for (int e = lane; e < 32 * kMlpF; e += 64) {
      const int r = e / kMlpF, k = e % kMlpF;
      xs[e] = (r0 + r < B) ? X[(size_t)idx[r0 + r] * kMlpF + k] : 0.0f;
    }
    __syncthreads();
    const float y = forward_tile(xs, 0, 32, wb, P, s);
    const bool valid = r0 + row < B;
    const float t = valid ? T[(size_t)(r0 + row) * kMlpO + kb] : 0.0f;
    const float err = valid ? (y - t) : 0.0f;
    s.dy[row][kb] = err * scale;
    atomicAdd(loss, err * err);
    __syncthreads();
    // layer 3 gradients
    for (int o = 0; o < kMlpO; ++o) {
      float g = 0.0f;
      for (int r = 0; r < 32; ++r) g += s.h2[r][lane] * s.dy[r][o];
      atomicAdd(&G[kMlpOffW3 + lane * kMlpO + o], g);
    }
    if (lane < kMlpO) {
      float g = 0.0f;
      for (int r = 0; r < 32; ++r) g += s.dy[r][lane];
      atomicAdd(&G[kMlpOffB3 + lane], g);
    }
    // dH2 = dY W3^T masked by ReLU
    for (int c = 32 * kb; c < 32 * kb + 32; ++c) {
      float g = 0.0f;
      for (int o = 0; o < kMlpO; ++o) g += s.dy[row][o] * from_bf16(wb[kMlpOffW3 + c * kMlpO + o]);
      g = s.h2[row][c] > 0.0f ? g : 0.0f;
      s.dh[row][c] = g;
      s.dhb[row][c] = to_bf16(g);
    }
    __syncthreads();
    {
      float g = 0.0f;
      for (int r = 0; r < 32; ++r) g += s.dh[r][lane];
      atomicAdd(&G[kMlpOffB2 + lane], g);
    }
    // dW2 = H1^T dH2: M = 64 (2 tiles), N = 64 (2 tiles), K = 32 rows (4 steps)
    for (int mt = 0; mt < 2; ++mt)
      for (int nt = 0; nt < 2; ++nt) {
        f32x16 acc = {};
        for (int st = 0; st < 4; ++st) {
          bf16x4 a, b;
          for (int i = 0; i < 4; ++i) {
            a[i] = (short)s.h1[8 * st + 4 * kb + i][mt * 32 + row];
            b[i] = (short)s.dhb[8 * st + 4 * kb + i][nt * 32 + row];
          }
          acc = mfma(a, b, acc);
        }
        for (int v = 0; v < 16; ++v)
          atomicAdd(&G[kMlpOffW2 + (mt * 32 + drow(v, lane)) * kMlpH + nt * 32 + row], acc[v]);
      }
    // dH1 = dH2 W2^T masked by ReLU: A = dH2 (32 x 64), B[k][n] = W2[n][k]
    f32x16 acc1[2] = {{}, {}};
    for (int st = 0; st < 8; ++st) {
      bf16x4 a;
      for (int i = 0; i < 4; ++i) a[i] = (short)s.dhb[row][8 * st + 4 * kb + i];
      for (int t = 0; t < 2; ++t) {
        bf16x4 b;
        for (int i = 0; i < 4; ++i) b[i] = (short)wb[kMlpOffW2 + (t * 32 + row) * kMlpH + 8 * st + 4 * kb + i];
        acc1[t] = mfma(a, b, acc1[t]);
      }
    }
    __syncthreads();
    for (int t = 0; t < 2; ++t)
      for (int v = 0; v < 16; ++v) {
        const int r = drow(v, lane), c = t * 32 + row;
        s.dh[r][c] = from_bf16(s.h1[r][c]) > 0.0f ? acc1[t][v] : 0.0f;
      }
    __syncthreads();
    // dW1 = X^T dH1 (8 x 64), db1
    {
      float gb = 0.0f;
      for (int r = 0; r < 32; ++r) gb += s.dh[r][lane];
      atomicAdd(&G[kMlpOffB1 + lane], gb);
      for (int k = 0; k < kMlpF; ++k) {
        float g = 0.0f;
        for (int r = 0; r < 32; ++r) g += from_bf16(to_bf16(s.x[r][k])) * s.dh[r][lane];
        atomicAdd(&G[kMlpOffW1 + k * kMlpH + lane], g);
      }
    }
    __syncthreads();
  }
}

// Adam on the FP32 master parameters, then the BF16 copy of the weights. The
// step number t >= 1 is read from device memory (so a captured graph can be
// replayed); bias corrections 1 - beta^t.
__global__ void mlp_adam_kernel(float* P, const float* G, float* M, float* V, uint16_t* wb, float lr, float b1, float b2,
                                float eps, const int* step) {
  const float t = (float)*step;
  const float c1 = 1.0f - powf(b1, t), c2 = 1.0f - powf(b2, t);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kMlpParams; i += gridDim.x * blockDim.x) {
    const float g = G[i];
    const float m = b1 * M[i] + (1.0f - b1) * g;
    const float v = b2 * V[i] + (1.0f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    const float p = P[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    P[i] = p;
    wb[i] = to_bf16(p);
  }
}

__global__ void step_bump_kernel(int* step) { *step += 1; }

__global__ void mlp_sync_kernel(const float* P, uint16_t* wb) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kMlpParams; i += gridDim.x * blockDim.x) wb[i] = to_bf16(P[i]);
}

// One 32x32x8 BF16 product, for checking the operand layout: A row-major 32x8,
// B row-major 8x32 (both already BF16), D row-major 32x32.
__global__ __launch_bounds__(64) void mfma_probe_kernel(const uint16_t* A, const uint16_t* B, float* D) {
  const int lane = threadIdx.x;
  bf16x4 a, b;
  for (int i = 0; i < 4; ++i) {
    a[i] = (short)A[(lane & 31) * 8 + 4 * (lane >> 5) + i];
    b[i] = (short)B[(4 * (lane >> 5) + i) * 32 + (lane & 31)];
  }
  f32x16 acc = {};
  acc = mfma(a, b, acc);
  for (int v = 0; v < 16; ++v) D[drow(v, lane) * 32 + (lane & 31)] = acc[v];
}

}  // namespace

hipError_t mlp_forward(const float* X, int N, const uint16_t* wb, const float* P, float* Y, hipStream_t s) {
  if (N <= 0) return hipSuccess;
  int blocks = (N + 31) / 32;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks), dim3(64), 0, s, X, N, wb, P, Y);
  return hipGetLastError();
}

hipError_t mlp_grad(const float* X, const float* T, const uint32_t* idx, int B, const uint16_t* wb, const float* P,
                    float* G, float* loss, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  int blocks = (B + 31) / 32;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(64), 0, s, X, T, idx, B, wb, P, G,
                     2.0f / (float)(B * kMlpO), loss);
  return hipGetLastError();
}

hipError_t mlp_adam(float* P, const float* G, float* M, float* V, uint16_t* wb, float lr, int* step, hipStream_t s) {
  // step counter first (t = previous + 1), then the update
  hipLaunchKernelGGL(step_bump_kernel, dim3(1), dim3(1), 0, s, step);
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, P, G, M, V, wb, lr, 0.9f,
                     0.999f, 1e-8f, (const int*)step);
  return hipGetLastError();
}

hipError_t mlp_sync(const float* P, uint16_t* wb, hipStream_t s) {
  hipLaunchKernelGGL(mlp_sync_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, P, wb);
  return hipGetLastError();
}

hipError_t mfma_probe(const uint16_t* A, const uint16_t* B, float* D, hipStream_t s) {
  hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, s, A, B, D);
  return hipGetLastError();
}

}  // namespace chl
