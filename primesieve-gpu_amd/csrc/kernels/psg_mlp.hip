// psg_mlp.hip -- FP32 cost model of the count path: a 6 -> 64 -> 64 -> 1
// multilayer perceptron (ReLU, ReLU, linear) that predicts log kernel time
// from (log2 start, log2 length, k, log2 segment bytes, waves per workgroup,
// log2 sieving primes).  Forward, backward and Adam are hand-written; the
// model only proposes tuning choices and never touches the counts.
//
// Parameters, one flat float array (PSG_MLP_PARAMS = 4673):
//   W1[64][6] @0, b1[64] @384, W2[64][64] @448 (W2[j][k]: unit j of layer 2
//   reads unit k of layer 1), b2[64] @4544, W3[64] @4608, b3 @4672.
//
// Layout of the work: one wave (64 lanes) per workgroup, lane j owns hidden
// unit j of both layers and keeps its weight rows in registers; a workgroup
// walks a tile of samples, exchanging the 64 activations of a layer through
// LDS.  FP32 throughout, FP32 accumulation.
#include <hip/hip_runtime.h>
#include <stdint.h>

#define H 64
#define NIN 6
#define OFF_W1 0
#define OFF_B1 384
#define OFF_W2 448
#define OFF_B2 4544
#define OFF_W3 4608
#define OFF_B3 4672
#define NPARAMS 4673
#define TILE 32

namespace {

__device__ __forceinline__ float wave_sum_f(float v)
{
  for (int d = 32; d >= 1; d >>= 1)
    v += __shfl_xor(v, d);
  return v;
}

}  // namespace

// y[s] for s in [0, n); also keeps the activations for the backward pass.
extern "C" __global__ void __launch_bounds__(64)
psg_mlp_forward(const float* __restrict__ P, const float* __restrict__ X, uint32_t n,
                float* __restrict__ H1, float* __restrict__ H2, float* __restrict__ Y)
{
  __shared__ float h1s[H];
  const uint32_t j = threadIdx.x;
  float w1[NIN], w2[H];
  for (int i = 0; i < NIN; i++) w1[i] = P[OFF_W1 + j * NIN + i];
  for (int k = 0; k < H; k++) w2[k] = P[OFF_W2 + j * H + k];
  const float b1 = P[OFF_B1 + j], b2 = P[OFF_B2 + j], w3 = P[OFF_W3 + j], b3 = P[OFF_B3];

  const uint32_t s0 = blockIdx.x * TILE;
  for (uint32_t s = s0; s < min(s0 + TILE, n); s++) {
    float a = b1;
    for (int i = 0; i < NIN; i++) a += w1[i] * X[s * NIN + i];
    const float h1 = fmaxf(a, 0.f);
    __syncthreads();
    h1s[j] = h1;
    __syncthreads();
    float c = b2;
    for (int k = 0; k < H; k++) c += w2[k] * h1s[k];
    const float h2 = fmaxf(c, 0.f);
    const float y = wave_sum_f(w3 * h2) + b3;
    if (H1) { H1[s * H + j] = h1; H2[s * H + j] = h2; }
    if (j == 0) Y[s] = y;
  }
}

// Gradients of L = (1/n) sum_s (y_s - t_s)^2, accumulated into G (zeroed by
// the caller) with float atomics; loss into *L.  Needs the forward
// activations H1, H2 and outputs Y.
extern "C" __global__ void __launch_bounds__(64)
psg_mlp_backward(const float* __restrict__ P, const float* __restrict__ X, const float* __restrict__ T,
                 const float* __restrict__ H1, const float* __restrict__ H2, const float* __restrict__ Y,
                 uint32_t n, float* __restrict__ G, float* __restrict__ L)
{
  __shared__ float g2s[H];
  __shared__ float h1s[H];
  const uint32_t j = threadIdx.x;
  float w2col[H];                       // column j of W2: W2[i][j] for all i
  for (int i = 0; i < H; i++) w2col[i] = P[OFF_W2 + i * H + j];
  const float w3 = P[OFF_W3 + j];

  float dW1[NIN] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float dW2[H];
  for (int k = 0; k < H; k++) dW2[k] = 0.f;
  float db1 = 0.f, db2 = 0.f, dW3 = 0.f, db3 = 0.f, loss = 0.f;
  const float scale = 2.f / (float) n;

  const uint32_t s0 = blockIdx.x * TILE;
  for (uint32_t s = s0; s < min(s0 + TILE, n); s++) {
    const float e = Y[s] - T[s];
    const float d = scale * e;             // dL/dy
    loss += e * e;
    const float h1 = H1[s * H + j];
    const float h2 = H2[s * H + j];
    dW3 += d * h2;
    db3 += d;
    const float g2 = (h2 > 0.f) ? d * w3 : 0.f;   // dL/d(pre-activation of unit j, layer 2)
    __syncthreads();
    g2s[j] = g2;
    h1s[j] = h1;
    __syncthreads();
    db2 += g2;
    for (int k = 0; k < H; k++) dW2[k] += g2 * h1s[k];
    float g1 = 0.f;                          // sum_i g2_i W2[i][j]
    for (int i = 0; i < H; i++) g1 += g2s[i] * w2col[i];
    g1 = (h1 > 0.f) ? g1 : 0.f;
    db1 += g1;
    for (int i = 0; i < NIN; i++) dW1[i] += g1 * X[s * NIN + i];
  }
  for (int i = 0; i < NIN; i++) atomicAdd(&G[OFF_W1 + j * NIN + i], dW1[i]);
  atomicAdd(&G[OFF_B1 + j], db1);
  for (int k = 0; k < H; k++) atomicAdd(&G[OFF_W2 + j * H + k], dW2[k]);
  atomicAdd(&G[OFF_B2 + j], db2);
  atomicAdd(&G[OFF_W3 + j], dW3);
  if (j == 0) {
    atomicAdd(&G[OFF_B3], db3);
    atomicAdd(L, loss / (float) n);
  }
}

// Adam: m <- b1 m + (1-b1) g; v <- b2 v + (1-b2) g^2;
// p <- p - lr * (m * c1) / (sqrt(v * c2) + eps), c1 = 1/(1-b1^t), c2 = 1/(1-b2^t).
extern "C" __global__ void __launch_bounds__(256)
psg_mlp_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
             float* __restrict__ V, uint32_t n, float lr, float b1, float b2, float eps,
             float c1, float c2)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float m = b1 * M[i] + (1.f - b1) * g;
  const float v = b2 * V[i] + (1.f - b2) * g * g;
  M[i] = m;
  V[i] = v;
  P[i] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
}

// zero n floats
extern "C" __global__ void __launch_bounds__(256)
psg_mlp_zero(float* __restrict__ A, uint32_t n)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[i] = 0.f;
}
