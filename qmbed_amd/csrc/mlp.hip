// FP32 multilayer perceptron of the learned Lanczos warm start: inputs are
// Hamiltonian couplings, outputs are coefficients over snapshot ground
// states. Shape in → h1 (tanh) → h2 (tanh) → out (linear), with
// in ≤ 8, h1, h2 ≤ 64, out ≤ 64. Parameters are one float array:
//   W1 [h1][in], b1 [h1], W2 [h2][h1], b2 [h2], W3 [out][h2], b3 [out].
// Loss: L = 1/(2n) Σ_s ‖y_s − t_s‖².
//
// Training runs in ONE workgroup for all epochs (full-batch Adam with an
// exponentially decaying learning rate): the weights live in LDS (twice:
// canonical for the backward pass, transposed for the forward pass), samples
// go through in chunks of 64, every thread owns a strided set of parameters
// whose gradient accumulators and Adam moments stay in registers, and sums
// over samples run in a fixed order, so training is deterministic. The
// network is tiny and latency-bound; one workgroup avoids a launch and a
// grid-wide reduction per epoch.
#include <hip/hip_runtime.h>

#include <cmath>

#include "internal.hpp"

namespace qmbed {
namespace {

constexpr int kTrainThreads = 1024;
constexpr int kChunk = 16;     // samples per forward/backward chunk
constexpr int kOwn = 9;        // parameters owned per thread (≥ 8896 / 1024)
constexpr int kMaxIn = 8, kMaxH = 64, kMaxOut = 64;

struct Shape {
  int in, h1, h2, out;
  __host__ __device__ int w1() const { return 0; }
  __host__ __device__ int b1() const { return h1 * in; }
  __host__ __device__ int w2() const { return b1() + h1; }
  __host__ __device__ int b2() const { return w2() + h2 * h1; }
  __host__ __device__ int w3() const { return b2() + h2; }
  __host__ __device__ int b3() const { return w3() + out * h2; }
  __host__ __device__ int count() const { return b3() + out; }
};

// forward pass of `ns` samples held in LDS (a0: [ns][in]); all threads of
// the block take part, output rows spread over (sample, neuron)
__device__ void forward_chunk(const Shape& sh, const float* P, const float* a0, float* a1, float* a2, float* y,
                              int ns) {
  for (int idx = threadIdx.x; idx < ns * sh.h1; idx += blockDim.x) {
    const int s = idx / sh.h1, j = idx % sh.h1;
    float acc = P[sh.b1() + j];
    for (int i = 0; i < sh.in; ++i) acc += P[sh.w1() + j * sh.in + i] * a0[s * kMaxIn + i];
    a1[s * kMaxH + j] = tanhf(acc);
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < ns * sh.h2; idx += blockDim.x) {
    const int s = idx / sh.h2, j = idx % sh.h2;
    float acc = P[sh.b2() + j];
    for (int i = 0; i < sh.h1; ++i) acc += P[sh.w2() + j * sh.h1 + i] * a1[s * kMaxH + i];
    a2[s * kMaxH + j] = tanhf(acc);
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < ns * sh.out; idx += blockDim.x) {
    const int s = idx / sh.out, o = idx % sh.out;
    float acc = P[sh.b3() + o];
    for (int i = 0; i < sh.h2; ++i) acc += P[sh.w3() + o * sh.h2 + i] * a2[s * kMaxH + i];
    y[s * kMaxOut + o] = acc;
  }
  __syncthreads();
}

// Inference: Y[s][o] for n samples; one block per kChunk samples.
__global__ __launch_bounds__(256) void k_mlp_forward(Shape sh, const float* __restrict__ Pg,
                                                     const float* __restrict__ X, int n, float* __restrict__ Y) {
  __shared__ float P[kMaxIn * kMaxH + kMaxH + kMaxH * kMaxH + kMaxH + kMaxOut * kMaxH + kMaxOut];
  __shared__ float a0[kChunk * kMaxIn], a1[kChunk * kMaxH], a2[kChunk * kMaxH], y[kChunk * kMaxOut];
  const int np = sh.count();
  for (int i = threadIdx.x; i < np; i += blockDim.x) P[i] = Pg[i];
  const int s0 = blockIdx.x * kChunk;
  const int ns = min(kChunk, n - s0);
  for (int idx = threadIdx.x; idx < ns * sh.in; idx += blockDim.x)
    a0[(idx / sh.in) * kMaxIn + idx % sh.in] = X[size_t(s0) * sh.in + idx];
  __syncthreads();
  forward_chunk(sh, P, a0, a1, a2, y, ns);
  for (int idx = threadIdx.x; idx < ns * sh.out; idx += blockDim.x)
    Y[size_t(s0) * sh.out + idx] = y[(idx / sh.out) * kMaxOut + idx % sh.out];
}

struct Adam {
  float lr, beta1, beta2, eps;
  float decay;  // the learning rate is multiplied by this after every epoch
};

constexpr int kTrainChunk = 64;  // samples per training chunk

// dynamic LDS of the training kernel, in floats: the parameters twice
// (canonical, read by the backward pass with consecutive lanes on
// consecutive columns, and with every weight matrix transposed, read the
// same way by the forward pass: no bank conflicts either way), then the
// activations and errors of one chunk
__host__ __device__ inline size_t train_lds_floats(int np) {
  return 2 * size_t(np) + size_t(kTrainChunk) * (kMaxIn + 4 * kMaxH + kMaxOut);
}

// position of parameter p in the transposed copy
__device__ inline int transposed(const Shape& sh, int p) {
  if (p < sh.b1()) {
    const int j = p / sh.in, i = p % sh.in;
    return sh.w1() + i * sh.h1 + j;
  }
  if (p >= sh.w2() && p < sh.b2()) {
    const int q = p - sh.w2(), j = q / sh.h1, i = q % sh.h1;
    return sh.w2() + i * sh.h2 + j;
  }
  if (p >= sh.w3() && p < sh.b3()) {
    const int q = p - sh.w3(), o = q / sh.h2, i = q % sh.h2;
    return sh.w3() + i * sh.out + o;
  }
  return p;  // biases
}

// Full-batch training (epochs > 0) or one gradient evaluation (epochs == 0:
// G and the loss are written, P is left unchanged). One block of
// kTrainThreads threads. loss_hist (optional) receives the loss of every
// epoch, evaluated before that epoch's update.
__global__ __launch_bounds__(kTrainThreads) void k_mlp_train(Shape sh, float* __restrict__ Pg,
                                                             const float* __restrict__ X,
                                                             const float* __restrict__ T, int n, int epochs,
                                                             Adam opt, float* __restrict__ G, float* loss_out,
                                                             float* loss_hist) {
  extern __shared__ float lds[];
  const int np = sh.count();
  float* P = lds;           // canonical
  float* Q = lds + np;      // transposed weights
  float* a0 = lds + 2 * np;  // [kTrainChunk][kMaxIn]
  float* a1 = a0 + kTrainChunk * kMaxIn;
  float* a2 = a1 + kTrainChunk * kMaxH;
  float* d1 = a2 + kTrainChunk * kMaxH;
  float* d2 = d1 + kTrainChunk * kMaxH;
  float* yd = d2 + kTrainChunk * kMaxH;  // outputs, then output errors
  __shared__ float red[kTrainThreads / 64];
  const int tid = threadIdx.x;
  for (int i = tid; i < np; i += blockDim.x) {
    const float v = Pg[i];
    P[i] = v;
    Q[transposed(sh, i)] = v;
  }
  float m[kOwn], v[kOwn];
#pragma unroll
  for (int k = 0; k < kOwn; ++k) m[k] = v[k] = 0.0f;
  const float inv_n = 1.0f / float(n);
  const int rounds = epochs > 0 ? epochs : 1;
  float b1t = 1.0f, b2t = 1.0f, lr = opt.lr;
  __syncthreads();
  for (int ep = 0; ep < rounds; ++ep) {
    float g[kOwn];
#pragma unroll
    for (int k = 0; k < kOwn; ++k) g[k] = 0.0f;
    float lsum = 0.0f;
    for (int s0 = 0; s0 < n; s0 += kTrainChunk) {
      const int ns = min(kTrainChunk, n - s0);
      for (int idx = tid; idx < ns * sh.in; idx += blockDim.x)
        a0[(idx / sh.in) * kMaxIn + idx % sh.in] = X[size_t(s0) * sh.in + idx];
      __syncthreads();
      // forward: consecutive lanes take consecutive neurons (transposed weights)
      for (int idx = tid; idx < ns * sh.h1; idx += blockDim.x) {
        const int s = idx / sh.h1, j = idx % sh.h1;
        float acc = Q[sh.b1() + j];
        for (int i = 0; i < sh.in; ++i) acc += Q[sh.w1() + i * sh.h1 + j] * a0[s * kMaxIn + i];
        a1[s * kMaxH + j] = tanhf(acc);
      }
      __syncthreads();
      for (int idx = tid; idx < ns * sh.h2; idx += blockDim.x) {
        const int s = idx / sh.h2, j = idx % sh.h2;
        float acc = Q[sh.b2() + j];
        for (int i = 0; i < sh.h1; ++i) acc += Q[sh.w2() + i * sh.h2 + j] * a1[s * kMaxH + i];
        a2[s * kMaxH + j] = tanhf(acc);
      }
      __syncthreads();
      // output and its error (loss 1/(2n) Σ ‖y − t‖²)
      for (int idx = tid; idx < ns * sh.out; idx += blockDim.x) {
        const int s = idx / sh.out, o = idx % sh.out;
        float acc = Q[sh.b3() + o];
        for (int i = 0; i < sh.h2; ++i) acc += Q[sh.w3() + i * sh.out + o] * a2[s * kMaxH + i];
        const float e = acc - T[size_t(s0 + s) * sh.out + o];
        yd[s * kMaxOut + o] = e * inv_n;
        lsum += 0.5f * e * e * inv_n;
      }
      __syncthreads();
      // backward (canonical weights: consecutive lanes on consecutive columns)
      for (int idx = tid; idx < ns * sh.h2; idx += blockDim.x) {
        const int s = idx / sh.h2, j = idx % sh.h2;
        float acc = 0.0f;
        for (int o = 0; o < sh.out; ++o) acc += P[sh.w3() + o * sh.h2 + j] * yd[s * kMaxOut + o];
        const float h = a2[s * kMaxH + j];
        d2[s * kMaxH + j] = acc * (1.0f - h * h);
      }
      __syncthreads();
      for (int idx = tid; idx < ns * sh.h1; idx += blockDim.x) {
        const int s = idx / sh.h1, j = idx % sh.h1;
        float acc = 0.0f;
        for (int k = 0; k < sh.h2; ++k) acc += P[sh.w2() + k * sh.h1 + j] * d2[s * kMaxH + k];
        const float h = a1[s * kMaxH + j];
        d1[s * kMaxH + j] = acc * (1.0f - h * h);
      }
      __syncthreads();
      // gradients of the owned parameters, summed over the chunk in order
#pragma unroll
      for (int k = 0; k < kOwn; ++k) {
        const int p = tid + k * kTrainThreads;
        if (p >= np) continue;
        float acc = 0.0f;
        if (p < sh.b1()) {
          const int j = p / sh.in, i = p % sh.in;
          for (int s = 0; s < ns; ++s) acc += d1[s * kMaxH + j] * a0[s * kMaxIn + i];
        } else if (p < sh.w2()) {
          const int j = p - sh.b1();
          for (int s = 0; s < ns; ++s) acc += d1[s * kMaxH + j];
        } else if (p < sh.b2()) {
          const int q = p - sh.w2(), j = q / sh.h1, i = q % sh.h1;
          for (int s = 0; s < ns; ++s) acc += d2[s * kMaxH + j] * a1[s * kMaxH + i];
        } else if (p < sh.w3()) {
          const int j = p - sh.b2();
          for (int s = 0; s < ns; ++s) acc += d2[s * kMaxH + j];
        } else if (p < sh.b3()) {
          const int q = p - sh.w3(), o = q / sh.h2, i = q % sh.h2;
          for (int s = 0; s < ns; ++s) acc += yd[s * kMaxOut + o] * a2[s * kMaxH + i];
        } else {
          const int o = p - sh.b3();
          for (int s = 0; s < ns; ++s) acc += yd[s * kMaxOut + o];
        }
        g[k] += acc;
      }
      __syncthreads();
    }
    // loss of this epoch (block sum in a fixed order)
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_down(lsum, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = lsum;
    __syncthreads();
    if (tid == 0) {
      float t = 0.0f;
      for (int w = 0; w < kTrainThreads / 64; ++w) t += red[w];
      if (loss_hist) loss_hist[ep] = t;
      if (loss_out) *loss_out = t;
    }
    if (epochs == 0) {
#pragma unroll
      for (int k = 0; k < kOwn; ++k) {
        const int p = tid + k * kTrainThreads;
        if (p < np) G[p] = g[k];
      }
      return;
    }
    // Adam step on the owned parameters (both LDS copies)
    b1t *= opt.beta1;
    b2t *= opt.beta2;
    const float c1 = 1.0f / (1.0f - b1t), c2 = 1.0f / (1.0f - b2t);
#pragma unroll
    for (int k = 0; k < kOwn; ++k) {
      const int p = tid + k * kTrainThreads;
      if (p >= np) continue;
      m[k] = opt.beta1 * m[k] + (1.0f - opt.beta1) * g[k];
      v[k] = opt.beta2 * v[k] + (1.0f - opt.beta2) * g[k] * g[k];
      const float nv = P[p] - lr * (m[k] * c1) / (sqrtf(v[k] * c2) + opt.eps);
      P[p] = nv;
      Q[transposed(sh, p)] = nv;
    }
    lr *= opt.decay;
    __syncthreads();
  }
  for (int i = tid; i < np; i += blockDim.x) Pg[i] = P[i];
}

Shape check_shape(int in, int h1, int h2, int out) {
  if (in < 1 || in > kMaxIn || h1 < 1 || h1 > kMaxH || h2 < 1 || h2 > kMaxH || out < 1 || out > kMaxOut)
    throw invalid_options("MLP shape must satisfy 1 ≤ in ≤ 8, 1 ≤ hidden ≤ 64, 1 ≤ out ≤ 64");
  return Shape{in, h1, h2, out};
}

}  // namespace

int mlp_parameter_count(int in, int h1, int h2, int out) { return check_shape(in, h1, h2, out).count(); }

void mlp_forward_device(int in, int h1, int h2, int out, const float* d_params, const float* d_x, int n, float* d_y,
                        hipStream_t s) {
  const Shape sh = check_shape(in, h1, h2, out);
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mlp_forward, dim3((n + kChunk - 1) / kChunk), dim3(256), 0, s, sh, d_params, d_x, n, d_y);
  QMBED_HIP(hipGetLastError());
}

void mlp_train_device(int in, int h1, int h2, int out, float* d_params, const float* d_x, const float* d_t, int n,
                      int epochs, float lr, float lr_decay, float* d_grad, float* d_loss, float* d_loss_hist,
                      hipStream_t s) {
  const Shape sh = check_shape(in, h1, h2, out);
  if (n <= 0) throw invalid_options("MLP training needs at least one sample");
  if (epochs < 0) throw invalid_options("MLP epochs must be non-negative");
  if (epochs == 0 && !d_grad) throw invalid_options("a gradient evaluation needs a gradient buffer");
  if (!(lr_decay > 0.0f && lr_decay <= 1.0f)) throw invalid_options("the learning-rate decay must lie in (0, 1]");
  const Adam opt{lr, 0.9f, 0.999f, 1.0e-8f, lr_decay};
  const size_t lds = train_lds_floats(sh.count()) * sizeof(float);
  QMBED_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mlp_train),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  hipLaunchKernelGGL(k_mlp_train, dim3(1), dim3(kTrainThreads), lds, s, sh, d_params, d_x, d_t, n, epochs, opt,
                     d_grad, d_loss, d_loss_hist);
  QMBED_HIP(hipGetLastError());
}

}  // namespace qmbed
