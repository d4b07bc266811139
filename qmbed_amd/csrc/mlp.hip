// FP32 multilayer perceptron of the learned Lanczos warm start: inputs are
// Hamiltonian couplings, outputs are coefficients over snapshot ground
// states. Shape in → h1 (tanh) → h2 (tanh) → out (linear), with
// in ≤ 8, h1, h2 ≤ 64, out ≤ 64. Parameters are one float array:
//   W1 [h1][in], b1 [h1], W2 [h2][h1], b2 [h2], W3 [out][h2], b3 [out].
// Loss: L = 1/(2n) Σ_s ‖y_s − t_s‖².
//
// Training runs in ONE workgroup for all epochs (full-batch Adam with an
// exponentially decaying learning rate), so an epoch needs no grid-wide
// reduction. The weights live in LDS, zero-padded to 64 × 64; samples go
// through in chunks of 64. The products of the hidden and output layers
// (forward, the backward errors and the two weight gradients, all 64 × 64 ×
// 64 with zero padding) run on the matrix cores: each of the 16 wavefronts
// owns one 16 × 16 output tile and issues v_mfma_f32_16x16x4_f32. The
// weight-gradient tiles stay in the owning lanes' accumulators for the whole
// epoch, and their Adam moments in the same lanes' registers; the input
// layer (K = in ≤ 8) and the biases run on the vector ALU. Sums over
// samples run in a fixed order, so training is deterministic.
#include <hip/hip_runtime.h>

#include <cmath>

#include "internal.hpp"

namespace qmbed {
namespace {

constexpr int kTrainThreads = 1024;  // 16 wavefronts: one 16 × 16 tile of a 64 × 64 product each
constexpr int kChunk = 16;           // samples per block of the inference kernel
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
constexpr int kLd = 68;          // row stride of the 64-wide LDS matrices (fewer bank conflicts)
constexpr int kMat = 64 * kLd;

// LDS of the training kernel, in floats: W1 [64][8], b1, b2, b3 [64] each,
// W2 and W3 [64][kLd], then a0 [64][8] and a1, a2, e, d2, d1 [64][kLd]
constexpr size_t kTrainLds = 64 * 8 + 3 * 64 + 2 * size_t(kMat) + 64 * 8 + 5 * size_t(kMat);

typedef float f32x4 __attribute__((ext_vector_type(4)));

// acc += A · B over K = 64 for the 16 × 16 tile (ti, tj) of the wavefront:
// A[i][k] = a[i·sai + k·sak], B[k][j] = b[k·sbk + j·sbj] (LDS). Operand
// lanes of v_mfma_f32_16x16x4_f32: A element (i = lane % 16, k = lane / 16),
// B element (k = lane / 16, j = lane % 16); the result element (i, j) sits in
// lane j + 16·(i / 4), register i % 4.
__device__ __forceinline__ f32x4 mfma_tile(f32x4 acc, const float* a, int sai, int sak, const float* b, int sbk,
                                           int sbj, int ti, int tj, int lane) {
  const int r = lane & 15, kq = lane >> 4;
  const float* pa = a + (ti * 16 + r) * sai + kq * sak;
  const float* pb = b + kq * sbk + (tj * 16 + r) * sbj;
#pragma unroll 4
  for (int k0 = 0; k0 < 64; k0 += 4)
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(pa[k0 * sak], pb[k0 * sbk], acc, 0, 0, 0);
  return acc;
}

__device__ __forceinline__ void adam_step(float& p, float g, float& m, float& v, const Adam& opt, float lr, float c1,
                                          float c2) {
  m = opt.beta1 * m + (1.0f - opt.beta1) * g;
  v = opt.beta2 * v + (1.0f - opt.beta2) * g * g;
  p -= lr * (m * c1) / (sqrtf(v * c2) + opt.eps);
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
  float* W1 = lds;            // [64][8]
  float* b1 = W1 + 64 * 8;    // [64]
  float* b2 = b1 + 64;
  float* b3 = b2 + 64;
  float* W2 = b3 + 64;        // [64][kLd]: W2[j][i], j < h2, i < h1
  float* W3 = W2 + kMat;      // W3[o][i], o < out, i < h2
  float* a0 = W3 + kMat;      // [64][8]
  float* a1 = a0 + 64 * 8;    // [64][kLd] per chunk: samples × neurons
  float* a2 = a1 + kMat;
  float* e = a2 + kMat;       // output errors (y − t)/n
  float* d2 = e + kMat;
  float* d1 = d2 + kMat;
  __shared__ float red[kTrainThreads / 64];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int ti = wave >> 2, tj = wave & 3;  // this wavefront's output tile
  // zero-padded parameters
  for (int i = tid; i < 64 * 8 + 3 * 64 + 2 * kMat; i += kTrainThreads) lds[i] = 0.0f;
  __syncthreads();
  for (int p = tid; p < sh.count(); p += kTrainThreads) {
    const float v = Pg[p];
    if (p < sh.b1()) W1[(p / sh.in) * 8 + p % sh.in] = v;
    else if (p < sh.w2()) b1[p - sh.b1()] = v;
    else if (p < sh.b2()) W2[((p - sh.w2()) / sh.h1) * kLd + (p - sh.w2()) % sh.h1] = v;
    else if (p < sh.w3()) b2[p - sh.b2()] = v;
    else if (p < sh.b3()) W3[((p - sh.w3()) / sh.h2) * kLd + (p - sh.w3()) % sh.h2] = v;
    else b3[p - sh.b3()] = v;
  }
  // parameters owned outside the matrix-core tiles: W1 (64 × 8) by threads
  // 0..511, b1, b2, b3 by threads 512..703
  const bool own_w1 = tid < 512;
  const int w1j = tid >> 3, w1i = tid & 7;
  const int bias = tid - 512;  // 0..191: b1, b2, b3
  const bool own_b = bias >= 0 && bias < 192;
  const bool live_w1 = own_w1 && w1j < sh.h1 && w1i < sh.in;
  const int bl = own_b ? bias / 64 : 0, bj = own_b ? bias % 64 : 0;
  const bool live_b = own_b && bj < (bl == 0 ? sh.h1 : bl == 1 ? sh.h2 : sh.out);
  float* bptr = own_b ? (bl == 0 ? b1 : bl == 1 ? b2 : b3) + bj : nullptr;
  float m_s = 0.0f, v_s = 0.0f;  // Adam moments of the owned W1 entry or bias
  // W2 and W3 tiles: lane holds rows 16·ti + 4·(lane/16) + r, column 16·tj + lane % 16
  float m2[4] = {0, 0, 0, 0}, v2[4] = {0, 0, 0, 0}, m3[4] = {0, 0, 0, 0}, v3[4] = {0, 0, 0, 0};
  const int trow = ti * 16 + 4 * (lane >> 4), tcol = tj * 16 + (lane & 15);
  const float inv_n = 1.0f / float(n);
  const int rounds = epochs > 0 ? epochs : 1;
  float b1t = 1.0f, b2t = 1.0f, lr = opt.lr;
  __syncthreads();
  for (int ep = 0; ep < rounds; ++ep) {
    f32x4 g2 = {0, 0, 0, 0}, g3 = {0, 0, 0, 0};
    float gs = 0.0f, lsum = 0.0f;
    for (int s0 = 0; s0 < n; s0 += kTrainChunk) {
      const int ns = min(kTrainChunk, n - s0);
      for (int idx = tid; idx < 64 * 8; idx += kTrainThreads) {
        const int s = idx >> 3, i = idx & 7;
        a0[idx] = (s < ns && i < sh.in) ? X[size_t(s0 + s) * sh.in + i] : 0.0f;
      }
      __syncthreads();
      // input layer on the vector ALU: a1 = tanh(b1 + W1 a0), 4 per thread
      for (int idx = tid; idx < 64 * 64; idx += kTrainThreads) {
        const int s = idx >> 6, j = idx & 63;
        float acc = b1[j];
#pragma unroll
        for (int i = 0; i < 8; ++i) acc += W1[j * 8 + i] * a0[s * 8 + i];
        a1[s * kLd + j] = tanhf(acc);
      }
      __syncthreads();
      // a2 = tanh(b2 + a1 W2ᵀ): A = a1 [s][i], B[k=i][j] = W2[j][i]
      {
        f32x4 z = mfma_tile(f32x4{0, 0, 0, 0}, a1, kLd, 1, W2, 1, kLd, ti, tj, lane);
#pragma unroll
        for (int r = 0; r < 4; ++r) a2[(trow + r) * kLd + tcol] = tanhf(z[r] + b2[tcol]);
      }
      __syncthreads();
      // y = b3 + a2 W3ᵀ and the output errors (zero beyond the chunk's samples
      // and the outputs, so padded rows and columns add nothing below)
      {
        f32x4 z = mfma_tile(f32x4{0, 0, 0, 0}, a2, kLd, 1, W3, 1, kLd, ti, tj, lane);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int s = trow + r, o = tcol;
          float err = 0.0f;
          if (s < ns && o < sh.out) {
            const float d = z[r] + b3[o] - T[size_t(s0 + s) * sh.out + o];
            err = d * inv_n;
            lsum += 0.5f * d * d * inv_n;
          }
          e[s * kLd + o] = err;
        }
      }
      __syncthreads();
      // gradient of W3 (A[m=o][k=s] = e[s][o], B[k=s][j=i] = a2[s][i]) and
      // d2 = (e W3) ⊙ (1 − a2²) (A = e [s][o], B[k=o][j=i] = W3[o][i])
      g3 = mfma_tile(g3, e, 1, kLd, a2, kLd, 1, ti, tj, lane);
      {
        f32x4 z = mfma_tile(f32x4{0, 0, 0, 0}, e, kLd, 1, W3, kLd, 1, ti, tj, lane);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float h = a2[(trow + r) * kLd + tcol];
          d2[(trow + r) * kLd + tcol] = z[r] * (1.0f - h * h);
        }
      }
      __syncthreads();
      // gradient of W2 (A[m=j][k=s] = d2[s][j], B[k=s][i] = a1[s][i]) and
      // d1 = (d2 W2) ⊙ (1 − a1²) (A = d2 [s][j], B[k=j][i] = W2[j][i])
      g2 = mfma_tile(g2, d2, 1, kLd, a1, kLd, 1, ti, tj, lane);
      {
        f32x4 z = mfma_tile(f32x4{0, 0, 0, 0}, d2, kLd, 1, W2, kLd, 1, ti, tj, lane);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float h = a1[(trow + r) * kLd + tcol];
          d1[(trow + r) * kLd + tcol] = z[r] * (1.0f - h * h);
        }
      }
      __syncthreads();
      // W1 and bias gradients on the vector ALU, summed over the chunk in order
      if (own_w1) {
        float acc = 0.0f;
        for (int s = 0; s < 64; ++s) acc += d1[s * kLd + w1j] * a0[s * 8 + w1i];
        gs += acc;
      } else if (own_b) {
        const float* src = bl == 0 ? d1 : bl == 1 ? d2 : e;
        float acc = 0.0f;
        for (int s = 0; s < 64; ++s) acc += src[s * kLd + bj];
        gs += acc;
      }
      __syncthreads();
    }
    // loss of this epoch (block sum in a fixed order)
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_down(lsum, o, 64);
    if (lane == 0) red[wave] = lsum;
    __syncthreads();
    if (tid == 0) {
      float t = 0.0f;
      for (int w = 0; w < kTrainThreads / 64; ++w) t += red[w];
      if (loss_hist) loss_hist[ep] = t;
      if (loss_out) *loss_out = t;
    }
    if (epochs == 0) {
      if (live_w1) G[w1j * sh.in + w1i] = gs;
      if (live_b) G[(bl == 0 ? sh.b1() : bl == 1 ? sh.b2() : sh.b3()) + bj] = gs;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (trow + r < sh.h2 && tcol < sh.h1) G[sh.w2() + (trow + r) * sh.h1 + tcol] = g2[r];
        if (trow + r < sh.out && tcol < sh.h2) G[sh.w3() + (trow + r) * sh.h2 + tcol] = g3[r];
      }
      return;
    }
    // Adam step on the owned parameters
    b1t *= opt.beta1;
    b2t *= opt.beta2;
    const float c1 = 1.0f / (1.0f - b1t), c2 = 1.0f / (1.0f - b2t);
    if (live_w1) adam_step(W1[w1j * 8 + w1i], gs, m_s, v_s, opt, lr, c1, c2);
    if (live_b) adam_step(*bptr, gs, m_s, v_s, opt, lr, c1, c2);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (trow + r < sh.h2 && tcol < sh.h1) adam_step(W2[(trow + r) * kLd + tcol], g2[r], m2[r], v2[r], opt, lr, c1, c2);
      if (trow + r < sh.out && tcol < sh.h2) adam_step(W3[(trow + r) * kLd + tcol], g3[r], m3[r], v3[r], opt, lr, c1, c2);
    }
    lr *= opt.decay;
    __syncthreads();
  }
  for (int p = tid; p < sh.count(); p += kTrainThreads) {
    float v;
    if (p < sh.b1()) v = W1[(p / sh.in) * 8 + p % sh.in];
    else if (p < sh.w2()) v = b1[p - sh.b1()];
    else if (p < sh.b2()) v = W2[((p - sh.w2()) / sh.h1) * kLd + (p - sh.w2()) % sh.h1];
    else if (p < sh.w3()) v = b2[p - sh.b2()];
    else if (p < sh.b3()) v = W3[((p - sh.w3()) / sh.h2) * kLd + (p - sh.w3()) % sh.h2];
    else v = b3[p - sh.b3()];
    Pg[p] = v;
  }
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
  const size_t lds = kTrainLds * sizeof(float);
  QMBED_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mlp_train),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  hipLaunchKernelGGL(k_mlp_train, dim3(1), dim3(kTrainThreads), lds, s, sh, d_params, d_x, d_t, n, epochs, opt,
                     d_grad, d_loss, d_loss_hist);
  QMBED_HIP(hipGetLastError());
}

}  // namespace qmbed
