// Lanczos vector kernels: dot/norm reductions, fused three-term updates,
// block Gram-Schmidt GEMVs (h = Vᵀw, w -= V h), Ritz assembly (V·C),
// precision conversion. Deterministic two-stage reductions (per-block
// partials, then one block sums them in a fixed order), no atomics.
//
// All vectors are HBM-resident; V is column-major with leading dimension n
// (one Krylov vector per column, contiguous), so every pass below streams
// whole columns with coalesced 64-lane accesses.
#include <algorithm>

#include "internal.hpp"

namespace qmbed {
namespace dev {

namespace {

constexpr int kT = 256;
constexpr int kMaxBlocks = 1024;
constexpr int kRowsPerBlockT = 2048;  // GEMV-T row chunk
constexpr int kColTile = 8;

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (blockDim.x + 63) / 64; ++w) r += sh[w];
  __syncthreads();
  return r;  // valid in thread 0
}

int nblocks(size_t n) { return int(std::min<size_t>(kMaxBlocks, std::max<size_t>(1, (n + kT * 4 - 1) / (kT * 4)))); }

template <int MODE>  // 0: x·y  1: x·x
__global__ void k_dot(const double* __restrict__ x, const double* __restrict__ y, size_t n, double* partial) {
  __shared__ double sh[8];
  double s = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT)
    s += MODE ? x[i] * x[i] : x[i] * y[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// complex: conj(x)·y -> (re, im) partials
template <int MODE>  // 0: conj(x)·y  1: |x|^2  2: max|x|
__global__ void k_zdot(const double2* __restrict__ x, const double2* __restrict__ y, size_t n, double* partial) {
  __shared__ double sh[8];
  double sr = 0, si = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double2 a = x[i];
    if (MODE == 0) {
      double2 b = y[i];
      sr += a.x * b.x + a.y * b.y;
      si += a.x * b.y - a.y * b.x;
    } else if (MODE == 1) {
      sr += a.x * a.x + a.y * a.y;
    } else {
      sr = fmax(sr, sqrt(a.x * a.x + a.y * a.y));
    }
  }
  if (MODE == 2) {
    for (int o = 32; o > 0; o >>= 1) sr = fmax(sr, __shfl_down(sr, o, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = sr;
    __syncthreads();
    if (threadIdx.x == 0) {
      double r = 0;
      for (int w = 0; w < kT / 64; ++w) r = fmax(r, sh[w]);
      partial[blockIdx.x] = r;
    }
    return;
  }
  sr = block_sum(sr, sh);
  if (MODE == 0) {
    double t = block_sum(si, sh);
    if (threadIdx.x == 0) partial[gridDim.x + blockIdx.x] = t;
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sr;
}

// out[c] = Σ_b partial[b * stride + c] (ordered), c < count; MODE 1: max
template <int MODE>
__global__ void k_reduce_cols(const double* partial, int nb, int stride, int count, double* out) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= count) return;
  double s = 0;
  for (int b = 0; b < nb; ++b) s = MODE ? fmax(s, partial[size_t(b) * stride + c]) : s + partial[size_t(b) * stride + c];
  out[c] = s;
}

// partial[rb * m + c] = Σ_{i in row block rb} V[c*n + i] * w[i]
__global__ void k_gemvt(const double* __restrict__ V, size_t n, int m, const double* __restrict__ w,
                        double* partial) {
  __shared__ double sh[kColTile][kT / 64];
  const size_t r0 = size_t(blockIdx.x) * kRowsPerBlockT;
  const int c0 = blockIdx.y * kColTile;
  const int nc = min(kColTile, m - c0);
  double acc[kColTile];
#pragma unroll
  for (int c = 0; c < kColTile; ++c) acc[c] = 0;
  const size_t r1 = min(n, r0 + kRowsPerBlockT);
  for (size_t i = r0 + threadIdx.x; i < r1; i += kT) {
    const double wi = w[i];
#pragma unroll
    for (int c = 0; c < kColTile; ++c)
      if (c < nc) acc[c] += V[size_t(c0 + c) * n + i] * wi;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < kColTile; ++c) {
    double v = acc[c];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) sh[c][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < nc) {
    double s = 0;
    for (int w2 = 0; w2 < kT / 64; ++w2) s += sh[threadIdx.x][w2];
    partial[size_t(blockIdx.x) * m + c0 + threadIdx.x] = s;
  }
}

// w -= V h  (m columns), partial[b] = Σ w_new^2 over the block's rows
__global__ void k_gemvn_sub(const double* __restrict__ V, size_t n, int m, const double* __restrict__ h, double* w,
                            double* partial) {
  __shared__ double sh_h[512];
  __shared__ double sh[8];
  for (int c = threadIdx.x; c < m; c += kT) sh_h[c] = h[c];
  __syncthreads();
  double s = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double acc = 0;
    int c = 0;
    for (; c + 4 <= m; c += 4) {
      acc += V[size_t(c) * n + i] * sh_h[c] + V[size_t(c + 1) * n + i] * sh_h[c + 1] +
             V[size_t(c + 2) * n + i] * sh_h[c + 2] + V[size_t(c + 3) * n + i] * sh_h[c + 3];
    }
    for (; c < m; ++c) acc += V[size_t(c) * n + i] * sh_h[c];
    double v = w[i] - acc;
    w[i] = v;
    s += v * v;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0 && partial) partial[blockIdx.x] = s;
}

// Y[:, q] = V C[:, q] for q < k (C is m×k column-major on the device)
template <int K>
__global__ void k_gemvn_multi(const double* __restrict__ V, size_t n, int m, const double* __restrict__ C, double* Y) {
  extern __shared__ double shc[];
  for (int i = threadIdx.x; i < m * K; i += kT) shc[i] = C[i];
  __syncthreads();
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double acc[K];
#pragma unroll
    for (int q = 0; q < K; ++q) acc[q] = 0;
    for (int c = 0; c < m; ++c) {
      const double v = V[size_t(c) * n + i];
#pragma unroll
      for (int q = 0; q < K; ++q) acc[q] += v * shc[q * m + c];
    }
#pragma unroll
    for (int q = 0; q < K; ++q) Y[size_t(q) * n + i] = acc[q];
  }
}

// w -= a v + b u with a, b read from device slots (b may be null); partial norm^2
__global__ void k_three_term(double* w, const double* __restrict__ v, const double* __restrict__ u,
                             const double* a_slot, const double* b_slot, size_t n, double* partial) {
  __shared__ double sh[8];
  const double a = *a_slot;
  const double b = b_slot ? *b_slot : 0.0;
  double s = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double x = w[i] - a * v[i];
    if (b_slot) x -= b * u[i];
    w[i] = x;
    s += x * x;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out = w / sqrt(nrm2_slot); optional FP32 copy
__global__ void k_normalize(const double* w, const double* nrm2_slot, size_t n, double* out, float* out32) {
  const double inv = 1.0 / sqrt(*nrm2_slot);
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double v = w[i] * inv;
    out[i] = v;
    if (out32) out32[i] = float(v);
  }
}

__global__ void k_fill(double* y, double v, size_t n) {
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) y[i] = v;
}
__global__ void k_axpy(double* y, double a, const double* x, size_t n) {
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) y[i] += a * x[i];
}
__global__ void k_scal(double* x, double a, size_t n) {
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) x[i] *= a;
}
__global__ void k_to_f32(float* y, const double* x, size_t n) {
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) y[i] = float(x[i]);
}
__global__ void k_to_f64(double* y, const float* x, size_t n) {
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) y[i] = double(x[i]);
}
__global__ void k_zaxpy(double2* y, double ar, double ai, const double2* x, size_t n) {
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double2 a = x[i], b = y[i];
    b.x += ar * a.x - ai * a.y;
    b.y += ar * a.y + ai * a.x;
    y[i] = b;
  }
}
__global__ void k_zscal(double2* x, double ar, double ai, size_t n) {
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double2 a = x[i];
    x[i] = make_double2(ar * a.x - ai * a.y, ar * a.y + ai * a.x);
  }
}

// complex GEMV-T: partial[rb*(2m) + 2c + {0,1}] = Σ conj(V[c]) w
__global__ void k_zgemvt(const double2* __restrict__ V, size_t n, int m, const double2* __restrict__ w,
                         double* partial) {
  __shared__ double sh[2 * kColTile][kT / 64];
  const size_t r0 = size_t(blockIdx.x) * kRowsPerBlockT;
  const int c0 = blockIdx.y * kColTile;
  const int nc = min(kColTile, m - c0);
  double ar[kColTile], ai[kColTile];
#pragma unroll
  for (int c = 0; c < kColTile; ++c) ar[c] = ai[c] = 0;
  const size_t r1 = min(n, r0 + kRowsPerBlockT);
  for (size_t i = r0 + threadIdx.x; i < r1; i += kT) {
    const double2 wi = w[i];
#pragma unroll
    for (int c = 0; c < kColTile; ++c)
      if (c < nc) {
        const double2 v = V[size_t(c0 + c) * n + i];
        ar[c] += v.x * wi.x + v.y * wi.y;
        ai[c] += v.x * wi.y - v.y * wi.x;
      }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < kColTile; ++c) {
    double v = ar[c], u = ai[c];
    for (int o = 32; o > 0; o >>= 1) {
      v += __shfl_down(v, o, 64);
      u += __shfl_down(u, o, 64);
    }
    if (lane == 0) {
      sh[2 * c][wid] = v;
      sh[2 * c + 1][wid] = u;
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * nc) {
    double s = 0;
    for (int w2 = 0; w2 < kT / 64; ++w2) s += sh[threadIdx.x][w2];
    partial[size_t(blockIdx.x) * 2 * m + 2 * c0 + threadIdx.x] = s;
  }
}

// w -= V h (complex)  / y = V c (complex; SUB=false overwrites y)
template <bool SUB>
__global__ void k_zgemvn(const double2* __restrict__ V, size_t n, int m, const double* __restrict__ h, double2* w) {
  __shared__ double sh_h[1024];
  for (int c = threadIdx.x; c < 2 * m; c += kT) sh_h[c] = h[c];
  __syncthreads();
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double ar = 0, ai = 0;
    for (int c = 0; c < m; ++c) {
      const double2 v = V[size_t(c) * n + i];
      const double hr = sh_h[2 * c], hi = sh_h[2 * c + 1];
      ar += v.x * hr - v.y * hi;
      ai += v.x * hi + v.y * hr;
    }
    if (SUB) {
      double2 x = w[i];
      w[i] = make_double2(x.x - ar, x.y - ai);
    } else {
      w[i] = make_double2(ar, ai);
    }
  }
}

// y = V c with V real (n×m), c complex
__global__ void k_dzgemvn(const double* __restrict__ V, size_t n, int m, const double* __restrict__ c, double2* y) {
  __shared__ double sh_c[1024];
  for (int j = threadIdx.x; j < 2 * m; j += kT) sh_c[j] = c[j];
  __syncthreads();
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double ar = 0, ai = 0;
    for (int j = 0; j < m; ++j) {
      const double v = V[size_t(j) * n + i];
      ar += v * sh_c[2 * j];
      ai += v * sh_c[2 * j + 1];
    }
    y[i] = make_double2(ar, ai);
  }
}

}  // namespace

Reducer::Reducer(size_t slots) : partial(std::max<size_t>(slots, 2 * kMaxBlocks * 2)), result(64), host(64) {}

void fill(double* y, double v, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(kT), 0, s, y, v, n);
}
void copy(double* y, const double* x, size_t n, hipStream_t s) {
  QMBED_HIP(hipMemcpyAsync(y, x, n * sizeof(double), hipMemcpyDeviceToDevice, s));
}
void axpy(double* y, double a, const double* x, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(nblocks(n)), dim3(kT), 0, s, y, a, x, n);
}
void scal(double* x, double a, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_scal, dim3(nblocks(n)), dim3(kT), 0, s, x, a, n);
}
void to_f32(float* y, const double* x, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_to_f32, dim3(nblocks(n)), dim3(kT), 0, s, y, x, n);
}
void to_f64(double* y, const float* x, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_to_f64, dim3(nblocks(n)), dim3(kT), 0, s, y, x, n);
}

static double reduce_to_host(Reducer& r, int nb, int mode, hipStream_t s) {
  if (mode == 0) hipLaunchKernelGGL(k_reduce_cols<0>, dim3(1), dim3(64), 0, s, r.partial.data(), nb, 1, 1, r.result.data());
  else hipLaunchKernelGGL(k_reduce_cols<1>, dim3(1), dim3(64), 0, s, r.partial.data(), nb, 1, 1, r.result.data());
  QMBED_HIP(hipMemcpyAsync(r.host.data(), r.result.data(), sizeof(double), hipMemcpyDeviceToHost, s));
  QMBED_HIP(hipStreamSynchronize(s));
  return r.host[0];
}

double dot(const double* x, const double* y, size_t n, hipStream_t s, Reducer& r) {
  int nb = nblocks(n);
  hipLaunchKernelGGL(k_dot<0>, dim3(nb), dim3(kT), 0, s, x, y, n, r.partial.data());
  return reduce_to_host(r, nb, 0, s);
}
double nrm2(const double* x, size_t n, hipStream_t s, Reducer& r) {
  int nb = nblocks(n);
  hipLaunchKernelGGL(k_dot<1>, dim3(nb), dim3(kT), 0, s, x, x, n, r.partial.data());
  return std::sqrt(reduce_to_host(r, nb, 0, s));
}

void gemv_t(const double* V, size_t n, int m, const double* w, double* d_h, hipStream_t s, Reducer& r) {
  int rb = int((n + kRowsPerBlockT - 1) / kRowsPerBlockT);
  if (size_t(rb) * m > r.partial.size()) r.partial.resize(size_t(rb) * m * 2);
  dim3 grid(rb, (m + kColTile - 1) / kColTile);
  hipLaunchKernelGGL(k_gemvt, grid, dim3(kT), 0, s, V, n, m, w, r.partial.data());
  hipLaunchKernelGGL(k_reduce_cols<0>, dim3((m + 63) / 64), dim3(64), 0, s, r.partial.data(), rb, m, m, d_h);
}

void gemv_n_sub(const double* V, size_t n, int m, const double* d_h, double* w, hipStream_t s) {
  hipLaunchKernelGGL(k_gemvn_sub, dim3(nblocks(n)), dim3(kT), 0, s, V, n, m, d_h, w, nullptr);
}

void gemv_n(const double* V, size_t n, int m, const double* d_c, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_gemvn_multi<1>, dim3(nblocks(n)), dim3(kT), m * sizeof(double), s, V, n, m, d_c, y);
}

void zaxpy(cplx* y, cplx a, const cplx* x, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_zaxpy, dim3(nblocks(n)), dim3(kT), 0, s, reinterpret_cast<double2*>(y), a.real(), a.imag(),
                     reinterpret_cast<const double2*>(x), n);
}
void zscal(cplx* x, cplx a, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_zscal, dim3(nblocks(n)), dim3(kT), 0, s, reinterpret_cast<double2*>(x), a.real(), a.imag(), n);
}
void zdscal(cplx* x, double a, size_t n, hipStream_t s) { zscal(x, cplx(a, 0), n, s); }

cplx zdot(const cplx* x, const cplx* y, size_t n, hipStream_t s, Reducer& r) {
  int nb = nblocks(n);
  hipLaunchKernelGGL(k_zdot<0>, dim3(nb), dim3(kT), 0, s, reinterpret_cast<const double2*>(x),
                     reinterpret_cast<const double2*>(y), n, r.partial.data());
  hipLaunchKernelGGL(k_reduce_cols<0>, dim3(1), dim3(64), 0, s, r.partial.data(), nb, 1, 1, r.result.data());
  hipLaunchKernelGGL(k_reduce_cols<0>, dim3(1), dim3(64), 0, s, r.partial.data() + nb, nb, 1, 1, r.result.data() + 1);
  QMBED_HIP(hipMemcpyAsync(r.host.data(), r.result.data(), 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  QMBED_HIP(hipStreamSynchronize(s));
  return cplx(r.host[0], r.host[1]);
}
double znrm2(const cplx* x, size_t n, hipStream_t s, Reducer& r) {
  int nb = nblocks(n);
  hipLaunchKernelGGL(k_zdot<1>, dim3(nb), dim3(kT), 0, s, reinterpret_cast<const double2*>(x),
                     reinterpret_cast<const double2*>(x), n, r.partial.data());
  return std::sqrt(reduce_to_host(r, nb, 0, s));
}
double zamax(const cplx* x, size_t n, hipStream_t s, Reducer& r) {
  int nb = nblocks(n);
  hipLaunchKernelGGL(k_zdot<2>, dim3(nb), dim3(kT), 0, s, reinterpret_cast<const double2*>(x),
                     reinterpret_cast<const double2*>(x), n, r.partial.data());
  return reduce_to_host(r, nb, 1, s);
}

void zgemv_t(const cplx* V, size_t n, int m, const cplx* w, cplx* d_h, hipStream_t s, Reducer& r) {
  int rb = int((n + kRowsPerBlockT - 1) / kRowsPerBlockT);
  if (size_t(rb) * 2 * m > r.partial.size()) r.partial.resize(size_t(rb) * 4 * m);
  dim3 grid(rb, (m + kColTile - 1) / kColTile);
  hipLaunchKernelGGL(k_zgemvt, grid, dim3(kT), 0, s, reinterpret_cast<const double2*>(V), n, m,
                     reinterpret_cast<const double2*>(w), r.partial.data());
  hipLaunchKernelGGL(k_reduce_cols<0>, dim3((2 * m + 63) / 64), dim3(64), 0, s, r.partial.data(), rb, 2 * m, 2 * m,
                     reinterpret_cast<double*>(d_h));
}
void zgemv_n_sub(const cplx* V, size_t n, int m, const cplx* d_h, cplx* w, hipStream_t s) {
  hipLaunchKernelGGL(k_zgemvn<true>, dim3(nblocks(n)), dim3(kT), 0, s, reinterpret_cast<const double2*>(V), n, m,
                     reinterpret_cast<const double*>(d_h), reinterpret_cast<double2*>(w));
}
void zgemv_n(const cplx* V, size_t n, int m, const cplx* d_c, cplx* y, hipStream_t s) {
  hipLaunchKernelGGL(k_zgemvn<false>, dim3(nblocks(n)), dim3(kT), 0, s, reinterpret_cast<const double2*>(V), n, m,
                     reinterpret_cast<const double*>(d_c), reinterpret_cast<double2*>(y));
}
void dzgemv_n(const double* V, size_t n, int m, const cplx* d_c, cplx* y, hipStream_t s) {
  hipLaunchKernelGGL(k_dzgemvn, dim3(nblocks(n)), dim3(kT), 0, s, V, n, m, reinterpret_cast<const double*>(d_c),
                     reinterpret_cast<double2*>(y));
}

void zdot_batch(const cplx* X, const cplx* Y, size_t n, int nvec, cplx* out, hipStream_t s, Reducer& r) {
  for (int c = 0; c < nvec; ++c) out[c] = zdot(X + size_t(c) * n, Y + size_t(c) * n, n, s, r);
}
void znrm2_batch(const cplx* X, size_t n, int nvec, double* out, hipStream_t s, Reducer& r) {
  for (int c = 0; c < nvec; ++c) out[c] = znrm2(X + size_t(c) * n, n, s, r);
}
void zamax_batch(const cplx* X, size_t n, int nvec, double* out, hipStream_t s, Reducer& r) {
  for (int c = 0; c < nvec; ++c) out[c] = zamax(X + size_t(c) * n, n, s, r);
}

// ---- device-scalar Lanczos helpers (no host round trip) ----
void lanczos_three_term(double* w, const double* v, const double* u, const double* a_slot, const double* b_slot,
                        size_t n, double* out_nrm2_slot, Reducer& r, hipStream_t s) {
  int nb = nblocks(n);
  hipLaunchKernelGGL(k_three_term, dim3(nb), dim3(kT), 0, s, w, v, u, a_slot, b_slot, n, r.partial.data());
  hipLaunchKernelGGL(k_reduce_cols<0>, dim3(1), dim3(64), 0, s, r.partial.data(), nb, 1, 1, out_nrm2_slot);
}
void dot_to_slot(const double* x, const double* y, size_t n, double* slot, Reducer& r, hipStream_t s) {
  int nb = nblocks(n);
  hipLaunchKernelGGL(k_dot<0>, dim3(nb), dim3(kT), 0, s, x, y, n, r.partial.data());
  hipLaunchKernelGGL(k_reduce_cols<0>, dim3(1), dim3(64), 0, s, r.partial.data(), nb, 1, 1, slot);
}
void gemv_n_sub_norm(const double* V, size_t n, int m, const double* d_h, double* w, double* nrm2_slot, Reducer& r,
                     hipStream_t s) {
  int nb = nblocks(n);
  hipLaunchKernelGGL(k_gemvn_sub, dim3(nb), dim3(kT), 0, s, V, n, m, d_h, w, r.partial.data());
  hipLaunchKernelGGL(k_reduce_cols<0>, dim3(1), dim3(64), 0, s, r.partial.data(), nb, 1, 1, nrm2_slot);
}
void normalize_by_slot(const double* w, const double* nrm2_slot, size_t n, double* out, float* out32, hipStream_t s) {
  hipLaunchKernelGGL(k_normalize, dim3(nblocks(n)), dim3(kT), 0, s, w, nrm2_slot, n, out, out32);
}
void gemm_vc(const double* V, size_t n, int m, const double* d_C, int k, double* Y, hipStream_t s) {
  // Y (n×k) = V (n×m) C (m×k), processed in groups of up to 8 columns
  for (int q0 = 0; q0 < k; q0 += 8) {
    int kk = std::min(8, k - q0);
    const double* C = d_C + size_t(q0) * m;
    double* Yq = Y + size_t(q0) * n;
    size_t sm = size_t(m) * kk * sizeof(double);
    switch (kk) {
      case 1: hipLaunchKernelGGL(k_gemvn_multi<1>, dim3(nblocks(n)), dim3(kT), sm, s, V, n, m, C, Yq); break;
      case 2: hipLaunchKernelGGL(k_gemvn_multi<2>, dim3(nblocks(n)), dim3(kT), sm, s, V, n, m, C, Yq); break;
      case 3: hipLaunchKernelGGL(k_gemvn_multi<3>, dim3(nblocks(n)), dim3(kT), sm, s, V, n, m, C, Yq); break;
      case 4: hipLaunchKernelGGL(k_gemvn_multi<4>, dim3(nblocks(n)), dim3(kT), sm, s, V, n, m, C, Yq); break;
      case 5: hipLaunchKernelGGL(k_gemvn_multi<5>, dim3(nblocks(n)), dim3(kT), sm, s, V, n, m, C, Yq); break;
      case 6: hipLaunchKernelGGL(k_gemvn_multi<6>, dim3(nblocks(n)), dim3(kT), sm, s, V, n, m, C, Yq); break;
      case 7: hipLaunchKernelGGL(k_gemvn_multi<7>, dim3(nblocks(n)), dim3(kT), sm, s, V, n, m, C, Yq); break;
      default: hipLaunchKernelGGL(k_gemvn_multi<8>, dim3(nblocks(n)), dim3(kT), sm, s, V, n, m, C, Yq); break;
    }
  }
}

}  // namespace dev
}  // namespace qmbed
