// Device-scalar Lanczos building blocks for real (double) and complex
// (double2) Krylov vectors. Scalars (α_j, β_j, norms, overlap vectors) stay in
// device slots so an iteration needs no host round trip except the DGKS
// decision; every pass streams whole columns of V (column-major, ld = n).
#include <algorithm>

#include "internal.hpp"
#include "lanczos_kernels.hpp"

namespace qmbed {
namespace lk {

namespace {

constexpr int kT = 256;
constexpr int kMaxBlocks = 1024;
constexpr int kRowsT = 2048;
constexpr int kTile = 8;

__device__ __forceinline__ double re_of(double a) { return a; }
__device__ __forceinline__ double re_of(double2 a) { return a.x; }
__device__ __forceinline__ double sq(double a) { return a * a; }
__device__ __forceinline__ double sq(double2 a) { return a.x * a.x + a.y * a.y; }
// conj(a) * b
__device__ __forceinline__ double cmul(double a, double b) { return a * b; }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
// a * b
__device__ __forceinline__ double mul(double a, double b) { return a * b; }
__device__ __forceinline__ double2 mul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double add(double a, double b) { return a + b; }
__device__ __forceinline__ double2 add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double sub(double a, double b) { return a - b; }
__device__ __forceinline__ double2 sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double scale(double a, double s) { return a * s; }
__device__ __forceinline__ double2 scale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
template <typename E>
__device__ __forceinline__ E zero();
template <>
__device__ __forceinline__ double zero<double>() {
  return 0.0;
}
template <>
__device__ __forceinline__ double2 zero<double2>() {
  return make_double2(0.0, 0.0);
}
__device__ __forceinline__ double shfl_down(double v, int o) { return __shfl_down(v, o, 64); }
__device__ __forceinline__ double2 shfl_down(double2 v, int o) {
  return make_double2(__shfl_down(v.x, o, 64), __shfl_down(v.y, o, 64));
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kT / 64; ++w) r += sh[w];
  __syncthreads();
  return r;
}

int nblocks(size_t n) { return int(std::min<size_t>(kMaxBlocks, std::max<size_t>(1, (n + kT * 4 - 1) / (kT * 4)))); }

// partial = Re Σ conj(x) y
template <typename E>
__global__ void k_redot(const E* __restrict__ x, const E* __restrict__ y, size_t n, double* partial) {
  __shared__ double sh[4];
  double s = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) s += re_of(cmul(x[i], y[i]));
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void k_sum(const double* partial, int nb, double* out, int mode) {
  if (threadIdx.x != 0) return;
  double s = 0;
  for (int b = 0; b < nb; ++b) s += partial[b];
  out[0] = mode ? sqrt(s) : s;
}

// w -= a v + b u  (a, b real device scalars), partial |w|^2
template <typename E>
__global__ void k_three(E* w, const E* __restrict__ v, const E* __restrict__ u, const double* a_slot,
                        const double* b_slot, size_t n, double* partial) {
  __shared__ double sh[4];
  const double a = *a_slot;
  const double b = b_slot ? *b_slot : 0.0;
  double s = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    E x = sub(w[i], scale(v[i], a));
    if (b_slot) x = sub(x, scale(u[i], b));
    w[i] = x;
    s += sq(x);
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// partial[rb*m + c] = Σ_rows conj(V_c) w
template <typename E>
__global__ void k_gemvt(const E* __restrict__ V, size_t n, int m, const E* __restrict__ w, E* partial) {
  __shared__ E sh[kTile][kT / 64];
  const size_t r0 = size_t(blockIdx.x) * kRowsT;
  const int c0 = blockIdx.y * kTile;
  const int nc = min(kTile, m - c0);
  E acc[kTile];
#pragma unroll
  for (int c = 0; c < kTile; ++c) acc[c] = zero<E>();
  const size_t r1 = min(n, r0 + kRowsT);
  for (size_t i = r0 + threadIdx.x; i < r1; i += kT) {
    const E wi = w[i];
#pragma unroll
    for (int c = 0; c < kTile; ++c)
      if (c < nc) acc[c] = add(acc[c], cmul(V[size_t(c0 + c) * n + i], wi));
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < kTile; ++c) {
    E v = acc[c];
    for (int o = 32; o > 0; o >>= 1) v = add(v, shfl_down(v, o));
    if (lane == 0) sh[c][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < nc) {
    E s = zero<E>();
    for (int w2 = 0; w2 < kT / 64; ++w2) s = add(s, sh[threadIdx.x][w2]);
    partial[size_t(blockIdx.x) * m + c0 + threadIdx.x] = s;
  }
}

template <typename E>
__global__ void k_reduce_cols(const E* partial, int nb, int m, E* out) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  E s = zero<E>();
  for (int b = 0; b < nb; ++b) s = add(s, partial[size_t(b) * m + c]);
  out[c] = s;
}

constexpr int kMaxCols = 512;

// w -= V h (columns [0, m)), partial |w|^2 ; h in device memory
template <typename E>
__global__ void k_gemvn_sub(const E* __restrict__ V, size_t n, int m, const E* __restrict__ h, E* w, double* partial) {
  __shared__ E sh_h[kMaxCols];
  __shared__ double sh[4];
  for (int c = threadIdx.x; c < m; c += kT) sh_h[c] = h[c];
  __syncthreads();
  double s = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    E acc = zero<E>();
    int c = 0;
    for (; c + 4 <= m; c += 4) {
      acc = add(acc, mul(V[size_t(c) * n + i], sh_h[c]));
      acc = add(acc, mul(V[size_t(c + 1) * n + i], sh_h[c + 1]));
      acc = add(acc, mul(V[size_t(c + 2) * n + i], sh_h[c + 2]));
      acc = add(acc, mul(V[size_t(c + 3) * n + i], sh_h[c + 3]));
    }
    for (; c < m; ++c) acc = add(acc, mul(V[size_t(c) * n + i], sh_h[c]));
    E v = sub(w[i], acc);
    w[i] = v;
    s += sq(v);
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0 && partial) partial[blockIdx.x] = s;
}

// out = w / sqrt(*nrm2); beta_slot (optional) = sqrt(*nrm2); optional FP32 copy
template <typename E, typename F>
__global__ void k_normalize(const E* w, const double* nrm2, size_t n, E* out, F* out32, double* beta_slot) {
  const double b = sqrt(*nrm2);
  const double inv = 1.0 / b;
  if (beta_slot && blockIdx.x == 0 && threadIdx.x == 0) *beta_slot = b;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    E v = scale(w[i], inv);
    out[i] = v;
    if (out32) {
      if constexpr (sizeof(E) == 8) out32[i] = float(re_of(v));
      else {
        out32[2 * i] = float(v.x);
        out32[2 * i + 1] = float(v.y);
      }
    }
  }
}

// Y[:, q] = V C[:, q], q < K; C column-major m×K with element type E
template <typename E, int K>
__global__ void k_gemm_vc(const E* __restrict__ V, size_t n, int m, const E* __restrict__ C, E* Y) {
  extern __shared__ double shc_raw[];
  E* shc = reinterpret_cast<E*>(shc_raw);
  for (int i = threadIdx.x; i < m * K; i += kT) shc[i] = C[i];
  __syncthreads();
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    E acc[K];
#pragma unroll
    for (int q = 0; q < K; ++q) acc[q] = zero<E>();
    for (int c = 0; c < m; ++c) {
      const E v = V[size_t(c) * n + i];
#pragma unroll
      for (int q = 0; q < K; ++q) acc[q] = add(acc[q], mul(v, shc[q * m + c]));
    }
#pragma unroll
    for (int q = 0; q < K; ++q) Y[size_t(q) * n + i] = acc[q];
  }
}

template <typename E>
__global__ void k_to32(const E* x, size_t n, float* y) {
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    if constexpr (sizeof(E) == 8) y[i] = float(re_of(x[i]));
    else {
      y[2 * i] = float(x[i].x);
      y[2 * i + 1] = float(x[i].y);
    }
  }
}
template <typename E>
__global__ void k_from32(const float* y, size_t n, E* x) {
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    if constexpr (sizeof(E) == 8) x[i] = double(y[i]);
    else x[i] = make_double2(double(y[2 * i]), double(y[2 * i + 1]));
  }
}

// r = y - lam * x ; partial |r|^2 (lam from device slot)
template <typename E>
__global__ void k_resid(const E* __restrict__ y, const E* __restrict__ x, const double* lam_slot, size_t n,
                        double* partial) {
  __shared__ double sh[4];
  const double lam = *lam_slot;
  double s = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT)
    s += sq(sub(y[i], scale(x[i], lam)));
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// deterministic_start (QMBED src/solve/mod.rs:729-742) with LCG jump-ahead:
// element i holds the (i+1)-th state of x <- a x + c from x0 = seed | 1.
__global__ void k_det_start(double* out, size_t n, uint64_t seed) {
  const uint64_t A = 6364136223846793005ULL, C = 1442695040888963407ULL;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    uint64_t steps = uint64_t(i) + 1;
    uint64_t ra = 1, rc = 0;  // accumulated affine map
    uint64_t pa = A, pc = C;  // current power
    while (steps) {
      if (steps & 1) {  // r = p ∘ r
        rc = pa * rc + pc;
        ra = pa * ra;
      }
      pc = pa * pc + pc;
      pa = pa * pa;
      steps >>= 1;
    }
    uint64_t x = ra * (seed | 1ULL) + rc;
    out[i] = double(x >> 11) / 9007199254740992.0 - 0.5;
  }
}

}  // namespace

Workspace::Workspace() : partial(64 * 1024), slots(4096) {}

template <typename E>
static const E* cE(const void* p) {
  return static_cast<const E*>(p);
}
template <typename E>
static E* mE(void* p) {
  return static_cast<E*>(p);
}

#define QMBED_DISPATCH(cplx_flag, ...)                                            \
  do {                                                                             \
    if (cplx_flag) {                                                               \
      using E = double2;                                                           \
      __VA_ARGS__;                                                                 \
    } else {                                                                       \
      using E = double;                                                            \
      __VA_ARGS__;                                                                 \
    }                                                                              \
  } while (0)

void redot_to_slot(bool c, const void* x, const void* y, size_t n, double* slot, Workspace& ws, hipStream_t s) {
  int nb = nblocks(n);
  QMBED_DISPATCH(c, hipLaunchKernelGGL(k_redot<E>, dim3(nb), dim3(kT), 0, s, cE<E>(x), cE<E>(y), n, ws.partial.data()));
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(64), 0, s, ws.partial.data(), nb, slot, 0);
}

void nrm2sq_to_slot(bool c, const void* x, size_t n, double* slot, Workspace& ws, hipStream_t s) {
  redot_to_slot(c, x, x, n, slot, ws, s);
}

void three_term(bool c, void* w, const void* v, const void* u, const double* a_slot, const double* b_slot, size_t n,
                double* nrm2_slot, Workspace& ws, hipStream_t s) {
  int nb = nblocks(n);
  QMBED_DISPATCH(c, hipLaunchKernelGGL(k_three<E>, dim3(nb), dim3(kT), 0, s, mE<E>(w), cE<E>(v), cE<E>(u), a_slot,
                                       b_slot, n, ws.partial.data()));
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(64), 0, s, ws.partial.data(), nb, nrm2_slot, 0);
}

void project_out(bool c, const void* V, size_t n, int m, void* w, void* d_h, double* nrm2_slot, Workspace& ws,
                 hipStream_t s) {
  // h = V^H w ; w -= V h ; nrm2 = |w|^2, in column chunks of kMaxCols
  const int rb = int((n + kRowsT - 1) / kRowsT);
  const size_t esz = c ? 16 : 8;
  size_t need = size_t(rb) * std::min(m, kMaxCols) * esz / 8;
  if (ws.partial.size() < need) ws.partial.resize(need * 2);
  for (int c0 = 0; c0 < m; c0 += kMaxCols) {
    int mc = std::min(kMaxCols, m - c0);
    bool last = c0 + mc >= m;
    QMBED_DISPATCH(c, {
      const E* Vc = cE<E>(V) + size_t(c0) * n;
      E* hp = mE<E>(d_h) + c0;
      E* part = reinterpret_cast<E*>(ws.partial.data());
      hipLaunchKernelGGL(k_gemvt<E>, dim3(rb, (mc + kTile - 1) / kTile), dim3(kT), 0, s, Vc, n, mc, cE<E>(w), part);
      hipLaunchKernelGGL(k_reduce_cols<E>, dim3((mc + 63) / 64), dim3(64), 0, s, part, rb, mc, hp);
    });
    (void)last;
  }
  // subtract after all overlaps are known (classical Gram-Schmidt pass)
  int nb = nblocks(n);
  DeviceArray<double>& part2 = ws.partial2;
  if (part2.size() < size_t(nb)) part2.resize(kMaxBlocks);
  for (int c0 = 0; c0 < m; c0 += kMaxCols) {
    int mc = std::min(kMaxCols, m - c0);
    bool last = c0 + mc >= m;
    QMBED_DISPATCH(c, hipLaunchKernelGGL(k_gemvn_sub<E>, dim3(nb), dim3(kT), 0, s, cE<E>(V) + size_t(c0) * n, n, mc,
                                         cE<E>(d_h) + c0, mE<E>(w), last ? part2.data() : nullptr));
  }
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(64), 0, s, part2.data(), nb, nrm2_slot, 0);
}

void normalize(bool c, const void* w, const double* nrm2_slot, size_t n, void* out, float* out32, double* beta_slot,
               hipStream_t s) {
  QMBED_DISPATCH(c, hipLaunchKernelGGL((k_normalize<E, float>), dim3(nblocks(n)), dim3(kT), 0, s, cE<E>(w), nrm2_slot,
                                       n, mE<E>(out), out32, beta_slot));
}

void gemm_vc(bool c, const void* V, size_t n, int m, const void* d_C, int k, void* Y, hipStream_t s) {
  const size_t esz = c ? 16 : 8;
  for (int q0 = 0; q0 < k; q0 += 4) {
    int kk = std::min(4, k - q0);
    size_t sm = size_t(m) * kk * esz;
    if (sm > 64 * 1024) throw unsupported_backend("Ritz assembly window exceeds the LDS coefficient budget");
    QMBED_DISPATCH(c, {
      const E* C = cE<E>(d_C) + size_t(q0) * m;
      E* Yq = mE<E>(Y) + size_t(q0) * n;
      switch (kk) {
        case 1: hipLaunchKernelGGL((k_gemm_vc<E, 1>), dim3(nblocks(n)), dim3(kT), sm, s, cE<E>(V), n, m, C, Yq); break;
        case 2: hipLaunchKernelGGL((k_gemm_vc<E, 2>), dim3(nblocks(n)), dim3(kT), sm, s, cE<E>(V), n, m, C, Yq); break;
        case 3: hipLaunchKernelGGL((k_gemm_vc<E, 3>), dim3(nblocks(n)), dim3(kT), sm, s, cE<E>(V), n, m, C, Yq); break;
        default: hipLaunchKernelGGL((k_gemm_vc<E, 4>), dim3(nblocks(n)), dim3(kT), sm, s, cE<E>(V), n, m, C, Yq); break;
      }
    });
  }
}

void to_f32(bool c, const void* x, size_t n, float* y, hipStream_t s) {
  QMBED_DISPATCH(c, hipLaunchKernelGGL(k_to32<E>, dim3(nblocks(n)), dim3(kT), 0, s, cE<E>(x), n, y));
}
void from_f32(bool c, const float* y, size_t n, void* x, hipStream_t s) {
  QMBED_DISPATCH(c, hipLaunchKernelGGL(k_from32<E>, dim3(nblocks(n)), dim3(kT), 0, s, y, n, mE<E>(x)));
}

void residual_to_slot(bool c, const void* y, const void* x, const double* lam_slot, size_t n, double* slot,
                      Workspace& ws, hipStream_t s) {
  int nb = nblocks(n);
  QMBED_DISPATCH(c, hipLaunchKernelGGL(k_resid<E>, dim3(nb), dim3(kT), 0, s, cE<E>(y), cE<E>(x), lam_slot, n,
                                       ws.partial.data()));
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(64), 0, s, ws.partial.data(), nb, slot, 1);
}

}  // namespace lk

void deterministic_start_device(double* d_out, size_t n, uint64_t seed, hipStream_t s) {
  hipLaunchKernelGGL(lk::k_det_start, dim3(lk::nblocks(n)), dim3(lk::kT), 0, s, d_out, n, seed);
  QMBED_HIP(hipGetLastError());
}

}  // namespace qmbed
