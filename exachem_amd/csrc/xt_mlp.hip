// Task cost model: a 7 -> 64 -> 64 -> 1 FP32 multilayer perceptron (ReLU
// hidden layers, linear output) with its inference, back-propagation and
// Adam kernels.  Every reduction over samples runs in a fixed order (one
// thread per parameter looping over samples), so a training run gives the
// same parameters bit for bit on every call.
//
// Parameter layout (kMlpParams floats):
//   W1[7][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3
// Loss: mean squared error over the batch.
#include <hip/hip_runtime.h>

#include "xt_device.h"

namespace xt {
namespace {

constexpr int IN = kMlpIn, H = kMlpHidden;
constexpr int OW1 = 0, OB1 = OW1 + IN * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H,
              OB3 = OW3 + H;
static_assert(OB3 + 1 == kMlpParams, "parameter layout");

// Forward pass, one thread per sample; parameters staged in LDS (all lanes
// read the same weight at a time).  A1/A2 (activations, n x H) are written
// when non-null.
__global__ void __launch_bounds__(256) k_mlp_forward(const float* __restrict__ P,
                                                     const float* __restrict__ X, long n,
                                                     float* __restrict__ Y, float* __restrict__ A1,
                                                     float* __restrict__ A2) {
  __shared__ float w[kMlpParams];
  for(int i = threadIdx.x; i < kMlpParams; i += blockDim.x) w[i] = P[i];
  __syncthreads();
  const long s = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if(s >= n) return;
  float x[IN];
#pragma unroll
  for(int i = 0; i < IN; i++) x[i] = X[s * IN + i];
  float a1[H];
#pragma unroll
  for(int h = 0; h < H; h++) {
    float z = w[OB1 + h];
#pragma unroll
    for(int i = 0; i < IN; i++) z = fmaf(x[i], w[OW1 + i * H + h], z);
    a1[h] = fmaxf(z, 0.f);
    if(A1) A1[s * H + h] = a1[h];
  }
  float y = w[OB3];
  for(int b = 0; b < H; b++) {
    float z = w[OB2 + b];
#pragma unroll
    for(int a = 0; a < H; a++) z = fmaf(a1[a], w[OW2 + a * H + b], z);
    const float a2 = fmaxf(z, 0.f);
    if(A2) A2[s * H + b] = a2;
    y = fmaf(a2, w[OW3 + b], y);
  }
  Y[s] = y;
}

// Output and hidden-layer error signals, one thread per sample:
//   DY = 2 (y - t) / n, D2 = DY w3 [a2 > 0], D1 = (W2 D2) [a1 > 0];
// SQ = (y - t)^2 for the loss.
__global__ void __launch_bounds__(256) k_mlp_backward(const float* __restrict__ P,
                                                      const float* __restrict__ Y,
                                                      const float* __restrict__ T, long n,
                                                      const float* __restrict__ A1,
                                                      const float* __restrict__ A2,
                                                      float* __restrict__ DY,
                                                      float* __restrict__ D1,
                                                      float* __restrict__ D2,
                                                      float* __restrict__ SQ) {
  __shared__ float w[kMlpParams];
  for(int i = threadIdx.x; i < kMlpParams; i += blockDim.x) w[i] = P[i];
  __syncthreads();
  const long s = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if(s >= n) return;
  const float r  = Y[s] - T[s];
  const float dy = 2.f * r / (float) n;
  SQ[s]          = r * r;
  DY[s]          = dy;
  float d2[H];
#pragma unroll
  for(int b = 0; b < H; b++) {
    d2[b]         = A2[s * H + b] > 0.f ? dy * w[OW3 + b] : 0.f;
    D2[s * H + b] = d2[b];
  }
  for(int a = 0; a < H; a++) {
    float g = 0.f;
#pragma unroll
    for(int b = 0; b < H; b++) g = fmaf(w[OW2 + a * H + b], d2[b], g);
    D1[s * H + a] = A1[s * H + a] > 0.f ? g : 0.f;
  }
}

// Parameter gradients, one thread per parameter, samples summed in order.
// Thread kMlpParams sums SQ into the mean loss.
__global__ void __launch_bounds__(256) k_mlp_grad(const float* __restrict__ X, long n,
                                                  const float* __restrict__ A1,
                                                  const float* __restrict__ A2,
                                                  const float* __restrict__ DY,
                                                  const float* __restrict__ D1,
                                                  const float* __restrict__ D2,
                                                  const float* __restrict__ SQ,
                                                  float* __restrict__ G, float* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if(p > kMlpParams) return;
  float g = 0.f;
  if(p < OB1) {
    const int i = p / H, h = p % H;
    for(long s = 0; s < n; s++) g = fmaf(X[s * IN + i], D1[s * H + h], g);
  }
  else if(p < OW2) {
    const int h = p - OB1;
    for(long s = 0; s < n; s++) g += D1[s * H + h];
  }
  else if(p < OB2) {
    const int a = (p - OW2) / H, b = (p - OW2) % H;
    for(long s = 0; s < n; s++) g = fmaf(A1[s * H + a], D2[s * H + b], g);
  }
  else if(p < OW3) {
    const int b = p - OB2;
    for(long s = 0; s < n; s++) g += D2[s * H + b];
  }
  else if(p < OB3) {
    const int b = p - OW3;
    for(long s = 0; s < n; s++) g = fmaf(DY[s], A2[s * H + b], g);
  }
  else if(p == OB3) {
    for(long s = 0; s < n; s++) g += DY[s];
  }
  else {  // p == kMlpParams: loss
    double l = 0.0;
    for(long s = 0; s < n; s++) l += SQ[s];
    loss[0] = (float) (l / (double) n);
    return;
  }
  G[p] = g;
}

// Adam update; the step count lives on the device (step[0]) so that a
// captured training step can be replayed.
__global__ void __launch_bounds__(256) k_mlp_adam(float* __restrict__ P, const float* __restrict__ G,
                                                  float* __restrict__ M1, float* __restrict__ M2,
                                                  const int* __restrict__ step, float lr) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if(p >= kMlpParams) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const int   t  = step[0] + 1;
  const float g  = G[p];
  const float m  = b1 * M1[p] + (1.f - b1) * g;
  const float v  = b2 * M2[p] + (1.f - b2) * g * g;
  M1[p]          = m;
  M2[p]          = v;
  const float mh = m / (1.f - powf(b1, (float) t));
  const float vh = v / (1.f - powf(b2, (float) t));
  P[p] -= lr * mh / (sqrtf(vh) + eps);
}

// step[0]: Adam steps taken so far; step[1]: value of step[0] when the
// current training call began (index base of the loss history).
__global__ void k_mlp_step_inc(int* step, const float* loss, float* loss_hist) {
  if(threadIdx.x == 0 && blockIdx.x == 0) {
    if(loss_hist) loss_hist[step[0] - step[1]] = loss[0];
    step[0] += 1;
  }
}

inline unsigned blocks(long n, int t) { return (unsigned) ((n + t - 1) / t); }

}  // namespace

int mlp_forward(const float* P, const float* X, long n, float* Y, float* A1, float* A2,
                void* stream) {
  if(n <= 0) return 0;
  hipLaunchKernelGGL(k_mlp_forward, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t) stream, P,
                     X, n, Y, A1, A2);
  return (int) hipGetLastError();
}

int mlp_loss_grad(const MlpWork& W, long n, void* stream) {
  hipStream_t st = (hipStream_t) stream;
  hipLaunchKernelGGL(k_mlp_forward, dim3(blocks(n, 256)), dim3(256), 0, st, W.P, W.X, n, W.Y, W.A1,
                     W.A2);
  hipLaunchKernelGGL(k_mlp_backward, dim3(blocks(n, 256)), dim3(256), 0, st, W.P, W.Y, W.T, n, W.A1,
                     W.A2, W.DY, W.D1, W.D2, W.SQ);
  hipLaunchKernelGGL(k_mlp_grad, dim3(blocks(kMlpParams + 1, 256)), dim3(256), 0, st, W.X, n, W.A1,
                     W.A2, W.DY, W.D1, W.D2, W.SQ, W.G, W.loss);
  return (int) hipGetLastError();
}

int mlp_train_step(const MlpWork& W, long n, float lr, float* loss_hist, void* stream) {
  hipStream_t st  = (hipStream_t) stream;
  int         err = mlp_loss_grad(W, n, stream);
  if(err) return err;
  hipLaunchKernelGGL(k_mlp_adam, dim3(blocks(kMlpParams, 256)), dim3(256), 0, st, W.P, W.G, W.M1,
                     W.M2, W.step, lr);
  hipLaunchKernelGGL(k_mlp_step_inc, dim3(1), dim3(64), 0, st, W.step, W.loss, loss_hist);
  return (int) hipGetLastError();
}

}  // namespace xt
