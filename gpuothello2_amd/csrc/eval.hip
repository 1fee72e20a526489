// Pattern evaluator kernels: feature indices, int16 evaluation, and the SGD
// forward/gradient and update kernels that train the evaluator (FP32 master
// weights, mean squared error against exact solver scores).
//
// The reference evaluates a float pattern model at think-mode leaves
// (eval.cu:123-142, thinker.cu:305) with weights it does not ship; here the
// model is trained by this project and used for move ordering in the exact
// solver (a wrong evaluation only changes node counts, never scores).
#include "bitboard.hpp"
#include "internal.hpp"
#include "pattern.hpp"

namespace go2 {

// base3[b] for b < 1024, built into LDS by each block.
__device__ __forceinline__ void fill_base3(uint16_t* t3) {
  for (int b = threadIdx.x; b < kBase3Size; b += blockDim.x) {
    uint32_t v = 0, p = 1;
    for (int k = 0; k < 10; ++k) {
      v += ((b >> k) & 1) * p;
      p *= 3;
    }
    t3[b] = (uint16_t)v;
  }
}

__global__ void features_kernel(const go2_board* boards, size_t n, uint32_t* feat, uint8_t* empt) {
  __shared__ uint16_t t3[kBase3Size];
  fill_base3(t3);
  __syncthreads();
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const go2_board b = boards[i];
  uint32_t* out = feat + i * kPatInstances;
  for_each_feature(b.player, b.opponent, t3, [&](int k, uint32_t f) { out[k] = f; });
  empt[i] = (uint8_t)(64 - popc64(b.player | b.opponent));
}

__global__ void eval_kernel(const int16_t* w, const int16_t* bias, const go2_board* boards, size_t n,
                            int32_t* out) {
  __shared__ uint16_t t3[kBase3Size];
  fill_base3(t3);
  __syncthreads();
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const go2_board b = boards[i];
  int32_t s = bias[64 - popc64(b.player | b.opponent)];
  for_each_feature(b.player, b.opponent, t3, [&](int, uint32_t f) { s += w[f]; });
  out[i] = s;
}

// Forward + gradient of the batch MSE: pred_i = sum_k w[f_ik] + bias[e_i],
// r_i = pred_i - y_i; every occurrence of a weight adds 2 r_i to its
// gradient and 1 to its count; the batch sum of r_i^2 goes to *sse.
__global__ void sgd_gradient_kernel(const float* w, const float* bias, const uint32_t* feat,
                                    const uint8_t* empt, const float* y, size_t n, float* grad,
                                    uint32_t* cnt, double* sse) {
  __shared__ float red[256];
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  float r2 = 0.f;
  if (i < n) {
    const uint32_t* f = feat + i * kPatInstances;
    float s = bias[empt[i]];
    for (int k = 0; k < kPatInstances; ++k) s += w[f[k]];
    const float r = s - y[i];
    r2 = r * r;
    const float g = 2.f * r;
    for (int k = 0; k < kPatInstances; ++k) {
      atomicAdd(&grad[f[k]], g);
      atomicAdd(&cnt[f[k]], 1u);
    }
    atomicAdd(&grad[kPatWeights + empt[i]], g);
    atomicAdd(&cnt[kPatWeights + empt[i]], 1u);
  }
  red[threadIdx.x] = r2;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(sse, (double)red[0]);
}

// Count-normalised step: w -= lr * grad / count for every weight seen in the
// batch; grad and count are cleared for the next batch.
__global__ void sgd_apply_kernel(float* w, float* bias, float* grad, uint32_t* cnt, float lr) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= kPatWeights + kPatBias) return;
  const uint32_t c = cnt[f];
  if (!c) return;
  const float step = lr * grad[f] / (float)c;
  if (f < kPatWeights) {
    w[f] -= step;
  } else {
    bias[f - kPatWeights] -= step;
  }
  grad[f] = 0.f;
  cnt[f] = 0u;
}

__global__ void quantize_kernel(const float* w, const float* bias, float scale, int16_t* w16,
                                int16_t* bias16) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= kPatWeights + kPatBias) return;
  const float v = (f < kPatWeights ? w[f] : bias[f - kPatWeights]) * scale;
  const float c = fminf(fmaxf(rintf(v), -32767.f), 32767.f);
  if (f < kPatWeights) {
    w16[f] = (int16_t)c;
  } else {
    bias16[f - kPatWeights] = (int16_t)c;
  }
}

static unsigned blocks_for(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace go2

using namespace go2;

extern "C" {

int go2_pattern_info(int* n_instances, int* n_weights, int* n_bias) {
  if (n_instances) *n_instances = kPatInstances;
  if (n_weights) *n_weights = kPatWeights;
  if (n_bias) *n_bias = kPatBias;
  return GO2_OK;
}

int go2_pattern_features_device(const go2_board* d_boards, size_t n, uint32_t* d_feat,
                                uint8_t* d_empties, void* stream) {
  if (n == 0) return GO2_OK;
  if (!d_boards || !d_feat || !d_empties) return fail(GO2_E_ARG, "null device array");
  hipLaunchKernelGGL(features_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     d_boards, n, d_feat, d_empties);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GO2_OK : hip_fail(e, "features kernel launch");
}

int go2_pattern_eval_device(const int16_t* d_w, const int16_t* d_bias, const go2_board* d_boards,
                            size_t n, int32_t* d_out, void* stream) {
  if (n == 0) return GO2_OK;
  if (!d_w || !d_bias || !d_boards || !d_out) return fail(GO2_E_ARG, "null device array");
  hipLaunchKernelGGL(eval_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, d_w,
                     d_bias, d_boards, n, d_out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GO2_OK : hip_fail(e, "eval kernel launch");
}

int go2_sgd_gradient_device(const float* d_w, const float* d_bias, const uint32_t* d_feat,
                            const uint8_t* d_empties, const float* d_y, size_t n, float* d_grad,
                            uint32_t* d_count, double* d_sse, void* stream) {
  if (n == 0) return GO2_OK;
  if (!d_w || !d_bias || !d_feat || !d_empties || !d_y || !d_grad || !d_count || !d_sse)
    return fail(GO2_E_ARG, "null device array");
  hipLaunchKernelGGL(sgd_gradient_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     d_w, d_bias, d_feat, d_empties, d_y, n, d_grad, d_count, d_sse);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GO2_OK : hip_fail(e, "sgd gradient kernel launch");
}

int go2_sgd_apply_device(float* d_w, float* d_bias, float* d_grad, uint32_t* d_count, float lr,
                         void* stream) {
  if (!d_w || !d_bias || !d_grad || !d_count) return fail(GO2_E_ARG, "null device array");
  hipLaunchKernelGGL(sgd_apply_kernel, dim3(blocks_for(kPatWeights + kPatBias, 256)), dim3(256), 0,
                     (hipStream_t)stream, d_w, d_bias, d_grad, d_count, lr);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GO2_OK : hip_fail(e, "sgd apply kernel launch");
}

int go2_quantize_weights_device(const float* d_w, const float* d_bias, float scale, int16_t* d_w16,
                                int16_t* d_bias16, void* stream) {
  if (!d_w || !d_bias || !d_w16 || !d_bias16) return fail(GO2_E_ARG, "null device array");
  hipLaunchKernelGGL(quantize_kernel, dim3(blocks_for(kPatWeights + kPatBias, 256)), dim3(256), 0,
                     (hipStream_t)stream, d_w, d_bias, scale, d_w16, d_bias16);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GO2_OK : hip_fail(e, "quantize kernel launch");
}

}  // extern "C"
